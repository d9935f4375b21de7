// C ABI of the fused time step (kernels in sles_rhs.h, instantiated per
// precision in sles_rhs_f64.hip / sles_rhs_f32.hip) and its kernel timing.
#include <mutex>

#include "sles_rhs.h"

// Workspace of sles_step_noalloc: the fused step's W fields and CFL partials
// for callers that cannot allocate device memory.  Each device that runs this
// library holds its own copy until the process ends; host code reaches it
// only through the address that step_scratch_init takes per device.
static __device__ __align__(256) unsigned char scratch[SLES_STEP_SCRATCH_BYTES];

namespace sles {
bool g_prof = false;
thread_local bool g_prof_off = false;
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_pool;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];
}  // namespace

// Device address of `scratch` on `device`, which must be the current device.
// Called once per device from sles_plan_create (the address differs per
// device); a failure is kept and reported by sles_step_noalloc.
void step_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
    (void)hipGetLastError();
    return;
  }
  g_scratch_ptr[device] = ptr;
  g_scratch_err[device].clear();
}

static void* step_scratch(int device, std::string* why) {
  if (device < 0 || device >= kMaxDevices) {
    *why = "device index out of range";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device])
    *why = g_scratch_err[device].empty() ? std::string("not initialised") : g_scratch_err[device];
  return g_scratch_ptr[device];
}
}  // namespace sles

namespace sles {
namespace {

// Workspace of the fused step: the W fields (planes padded by tune_pad) and
// the CFL partials; with padded planes also the drivers' copies of the RK
// state in the same layout (U_hat, U_hat0, U_hat1, aux: 12 fields, at
// mirror_offset), preceded by one 256-byte slot per member for the steps
// count at the start of a call.
size_t work_fields_bytes(const sles_plan_s* p, int les_model) {
  const size_t nw = les_model ? 9 : 6;
  return nw * (size_t)cplx_w(p) * (size_t)cplx_bytes(p) + sizeof(double) * ((size_t)p->n * p->n + 64);
}
size_t steps_slot_offset(const sles_plan_s* p, int les_model) {
  return (work_fields_bytes(p, les_model) + 255) / 256 * 256;
}
size_t mirror_offset(const sles_plan_s* p, int les_model) { return steps_slot_offset(p, les_model) + 256; }
constexpr int kMirrorFields = 12;

// Copies 3 spectral fields of N kx planes between the caller's arrays
// (reference layout: planes N (N/2+1) values apart, rows of N/2+1; members
// ref_m bytes apart) and the workspace's copies (planes pw values apart,
// split arrangement: rows of N/2, then the Nyquist column; members wrk_m
// bytes apart); fields lie N planes apart in both.  to_split: caller ->
// workspace, else workspace -> caller.  blockIdx.z is the member.  With
// `steps` set, member m copies only if its driver state counts a step taken
// since `before` saved the count (so a call that took no step leaves the
// caller's array as it was).
template <class E>
__global__ void k_relayout(E* ref, long ref_m, E* wrk, long pw, long wrk_m, int n, int to_split, const double* steps,
                           const double* before, long drive_stride, long before_stride) {
  const long m = blockIdx.z;
  if (steps) {
    const double now = *(const double*)((const char*)steps + m * drive_stride);
    const double then = *(const double*)((const char*)before + m * before_stride);
    if (now == then) return;
  }
  const long plane = blockIdx.y;  // field * n + kx plane
  const int nr = n / 2, nh = n / 2 + 1;
  const long main = (long)n * nr, total = (long)n * nh;
  E* R = (E*)((char*)ref + m * ref_m) + plane * total;
  E* Wk = (E*)((char*)wrk + m * wrk_m) + plane * pw;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
    // j: column of the split plane; r: the same value in the caller's plane
    const long r = j < main ? (j / nr) * nh + j % nr : (j - main) * nh + nr;
    if (to_split)
      Wk[j] = R[r];
    else
      R[r] = Wk[j];
  }
}

// before[m] = the steps count of member m's driver state, m < members
__global__ void k_save_steps(const double* steps, long drive_stride, double* before, long before_stride,
                             int members) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < members)
    *(double*)((char*)before + m * before_stride) = *(const double*)((const char*)steps + m * drive_stride);
}

// k_relayout over `members` members of one array (3 fields).
int relayout(const sles_plan_s* p, int to_split, void* ref, long ref_m, void* wrk, long wrk_m, int members,
             const double* steps, const double* before, long drive_stride, long before_stride, hipStream_t s) {
  const long per_plane = ncplx(p) / p->n;
  const dim3 grid((unsigned)((per_plane + 1023) / 1024), (unsigned)(3 * p->n), (unsigned)members);
  const long pw = plane_w(p);
  if (p->precision == SLES_FP64)
    hipLaunchKernelGGL(k_relayout<uint4>, grid, dim3(256), 0, s, (uint4*)ref, ref_m, (uint4*)wrk, pw, wrk_m, p->n,
                       to_split, steps, before, drive_stride, before_stride);
  else
    hipLaunchKernelGGL(k_relayout<uint2>, grid, dim3(256), 0, s, (uint2*)ref, ref_m, (uint2*)wrk, pw, wrk_m, p->n,
                       to_split, steps, before, drive_stride, before_stride);
  return check_launch("state relayout");
}

// The RK-state arrays a driver call works on: the caller's (reference
// layout) or, with padded workspace planes, copies in the workspace's layout
// that the call fills from the caller's arrays on entry and copies back on
// exit.
struct DriveState {
  void *uh, *uh0, *uh1, *aux;
  long ps;          // kx-plane stride (elements)
  int split;        // arrangement: 0 the caller's, 1 the workspace's
  long mstride;     // bytes between members' arrays
};

// Driver call entry: with padded workspace planes, point the call at the
// state copies in each member's workspace (members `wstride` bytes apart)
// and fill them from the caller's arrays (members `sbytes` apart): U_hat,
// and aux for ETD2 (which reads F_{n-1}); U_hat0 / U_hat1 / aux of the RK
// schemes are written by every step before they are read.  Also saves each
// member's steps count, which drive_exit compares.
int drive_enter(const sles_plan_s* p, int scheme, int les, void* uh, void* uh0, void* uh1, void* aux, int members,
                long sbytes, double* state, long dstride, void* work, long wstride, hipStream_t s, DriveState* ds) {
  const long dense = ncplx(p) / p->n;
  if (p->tune_pad == 0) {
    *ds = DriveState{uh, uh0, uh1, aux, dense, 0, sbytes};
    return 0;
  }
  // copies of the arrays the scheme touches: U_hat0 (RK4, RK3), U_hat1 (RK4), aux
  const long fbytes = cplx_w(p) * cplx_bytes(p);
  char* m = (char*)work + mirror_offset(p, les);
  const bool rk = scheme == SLES_SCHEME_RK4 || scheme == SLES_SCHEME_RK3;
  const bool rk4 = scheme == SLES_SCHEME_RK4;
  *ds = DriveState{m, rk && uh0 ? m + 3 * fbytes : nullptr, rk4 && uh1 ? m + 6 * fbytes : nullptr,
                   aux ? m + 9 * fbytes : nullptr, plane_w(p), 1, wstride};
  double* before = (double*)((char*)work + steps_slot_offset(p, les));
  hipLaunchKernelGGL(k_save_steps, dim3((unsigned)((members + 255) / 256)), dim3(256), 0, s,
                     state + SLES_DRIVE_STATE_STEPS, dstride, before, wstride, members);
  if (check_launch("driver steps count")) return -1;
  if (relayout(p, 1, uh, sbytes, ds->uh, wstride, members, nullptr, nullptr, 0, 0, s)) return -1;
  if (scheme == SLES_SCHEME_ETD2 && relayout(p, 1, aux, sbytes, ds->aux, wstride, members, nullptr, nullptr, 0, 0, s))
    return -1;
  return 0;
}

// Driver call exit: copy the state copies back to the caller's arrays --
// U_hat always, and ETD2's aux; U_hat0, U_hat1 and RK4's aux (those the
// scheme writes) only for members that took a step in this call (else the
// copies hold nothing of it).
int drive_exit(const sles_plan_s* p, int scheme, int les, void* uh, void* uh0, void* uh1, void* aux, int members,
               long sbytes, double* state, long dstride, void* work, long wstride, hipStream_t s,
               const DriveState& ds) {
  if (p->tune_pad == 0) return 0;
  const double* steps = state + SLES_DRIVE_STATE_STEPS;
  const double* before = (const double*)((char*)work + steps_slot_offset(p, les));
  if (relayout(p, 0, uh, sbytes, ds.uh, wstride, members, nullptr, nullptr, 0, 0, s)) return -1;
  if (ds.uh0 && relayout(p, 0, uh0, sbytes, ds.uh0, wstride, members, steps, before, dstride, wstride, s)) return -1;
  if (ds.uh1 && relayout(p, 0, uh1, sbytes, ds.uh1, wstride, members, steps, before, dstride, wstride, s)) return -1;
  if (ds.aux) {
    const bool always = scheme == SLES_SCHEME_ETD2;
    if (relayout(p, 0, aux, sbytes, ds.aux, wstride, members, always ? nullptr : steps, always ? nullptr : before,
                 dstride, wstride, s))
      return -1;
  }
  return 0;
}

}  // namespace
}  // namespace sles

using namespace sles;

extern "C" {

size_t sles_step_work_bytes(sles_plan p, int les_model) {
  if (!p) return 0;
  if (p->tune_pad == 0) return work_fields_bytes(p, les_model);
  return mirror_offset(p, les_model) + (size_t)kMirrorFields * (size_t)cplx_w(p) * (size_t)cplx_bytes(p);
}

int sles_step(sles_plan p, int scheme, const sles_step_params* prm, void* u_hat, void* u_hat0, void* u_hat1,
              void* aux, void* u_out, void* curl_out, double* scalars, void* work, void* stream) {
  if (!p || !prm || !u_hat || !work) return fail("sles_step: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_step: les_model must be 0, 1 or 2");
  if ((scheme == SLES_SCHEME_RK4 || scheme == SLES_SCHEME_RK3) && !u_hat0)
    return fail("sles_step: RK schemes need u_hat0");
  if (scheme == SLES_SCHEME_RK4 && !u_hat1) return fail("sles_step: RK4 needs u_hat1");
  if ((scheme == SLES_SCHEME_ETD2 || scheme == SLES_SCHEME_RHS || scheme == SLES_SCHEME_NONLINEAR) && !aux)
    return fail("sles_step: this scheme needs aux");
  if (prm->dt_device && !scalars) return fail("sles_step: dt_device needs scalars");
  if (prm->dt_device && (scheme == SLES_SCHEME_RHS || scheme == SLES_SCHEME_NONLINEAR))
    return fail("sles_step: dt_device only applies to time steps");
  hipStream_t s = (hipStream_t)stream;
  if (!sles_supported_n(p->n)) return fail("sles_step: unsupported grid size");
  const long ps = ncplx(p) / p->n;  // the caller's arrays: reference layout
  return p->precision == SLES_FP64
             ? step_f64(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, ps, 0, u_out, curl_out, scalars, work, s, nullptr)
             : step_f32(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, ps, 0, u_out, curl_out, scalars, work, s,
                        nullptr);
}

int sles_step_noalloc(sles_plan p, int scheme, const sles_step_params* prm, void* u_hat, void* u_hat0,
                      void* u_hat1, void* aux, void* u_out, void* curl_out, double* scalars, void* stream) {
  if (!p || !prm) return fail("sles_step_noalloc: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_step_noalloc: les_model must be 0, 1 or 2");
  // a single step uses the W fields and the CFL partials only (the drivers'
  // copies of the RK state that sles_step_work_bytes also counts are unused)
  const size_t need = sles::work_fields_bytes(p, prm->les_model);
  if (need > (size_t)SLES_STEP_SCRATCH_BYTES)
    return fail("sles_step_noalloc: the step needs " + std::to_string(need) + " bytes of workspace, above the " +
                std::to_string((size_t)SLES_STEP_SCRATCH_BYTES) + "-byte device scratch (N = " +
                std::to_string(p->n) + "); use sles_step with a caller workspace");
  std::string why;
  void* work = step_scratch(p->device, &why);
  if (!work) return fail("sles_step_noalloc: device scratch unavailable on device " + std::to_string(p->device) +
                         " (" + why + ")");
  // no kernel-timing events here: sles_profile_enable would create them
  g_prof_off = true;
  const int rc = sles_step(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, u_out, curl_out, scalars, work, stream);
  g_prof_off = false;
  return rc;
}

int sles_step_drive(sles_plan p, int scheme, const sles_step_params* prm, double cfl_first, int nsteps, void* u_hat,
                    void* u_hat0, void* u_hat1, void* aux, void* u_out, void* curl_out, double* state, void* work,
                    void* stream) {
  if (!p || !prm || !u_hat || !state || !work) return fail("sles_step_drive: null argument");
  if (scheme != SLES_SCHEME_RK4 && scheme != SLES_SCHEME_RK3 && scheme != SLES_SCHEME_ETD2)
    return fail("sles_step_drive: scheme must be RK4, RK3 or ETD2");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_step_drive: les_model must be 0, 1 or 2");
  if (scheme != SLES_SCHEME_ETD2 && !u_hat0) return fail("sles_step_drive: RK schemes need u_hat0");
  if (scheme == SLES_SCHEME_RK4 && !u_hat1) return fail("sles_step_drive: RK4 needs u_hat1");
  if (scheme == SLES_SCHEME_ETD2 && !aux) return fail("sles_step_drive: ETD2 needs aux");
  if (nsteps < 0) return fail("sles_step_drive: nsteps must be >= 0");
  if (!sles_supported_n(p->n)) return fail("sles_step_drive: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  if (nsteps == 0) return 0;
  DriveArgs drv{state, cfl_first};
  const long sbytes = 3 * ncplx(p) * cplx_bytes(p);
  DriveState ds{};
  if (drive_enter(p, scheme, prm->les_model, u_hat, u_hat0, u_hat1, aux, 1, sbytes, state, 0, work, 0, s, &ds))
    return -1;
  for (int i = 0; i < nsteps; ++i) {
    // U and curl are outputs only (each step recomputes the physical fields
    // it needs from U_hat, and dt from the z pass's max u): the steps before
    // the last would write values the next step overwrites unread
    const bool last = i + 1 == nsteps;
    void* uo = last ? u_out : nullptr;
    void* co = last ? curl_out : nullptr;
    drv.xinv_in = i > 0;  // read only where the step fuses (RK4, fuse_x)
    drv.xinv_out = !last;
    const int rc =
        p->precision == SLES_FP64
            ? step_f64(p, scheme, prm, ds.uh, ds.uh0, ds.uh1, ds.aux, ds.ps, ds.split, uo, co, nullptr, work, s,
                       nullptr, &drv)
            : step_f32(p, scheme, prm, ds.uh, ds.uh0, ds.uh1, ds.aux, ds.ps, ds.split, uo, co, nullptr, work, s,
                       nullptr, &drv);
    if (rc) return rc;
  }
  return drive_exit(p, scheme, prm->les_model, u_hat, u_hat0, u_hat1, aux, 1, sbytes, state, 0, work, 0, s, ds);
}

// Members of a batch lie at fixed strides: state fields packed (3 C each),
// workspaces rounded up to 256 bytes, driver states SLES_DRIVE_STATE_SIZE
// doubles.
static size_t batch_work_stride(sles_plan p, int les_model) {
  return (sles_step_work_bytes(p, les_model) + 255) / 256 * 256;
}

size_t sles_step_batch_work_bytes(sles_plan p, int les_model, int members) {
  if (!p || members < 1) return 0;
  return batch_work_stride(p, les_model) * (size_t)members;
}

int sles_step_drive_batch(sles_plan p, int scheme, const sles_step_params* prm, int members, const double* sgs_coeff,
                          double cfl_first, int nsteps, void* u_hat, void* u_hat0, void* u_hat1, void* aux,
                          double* state, void* work, void* stream) {
  if (!p || !prm || !u_hat || !state || !work) return fail("sles_step_drive_batch: null argument");
  if (scheme != SLES_SCHEME_RK4 && scheme != SLES_SCHEME_RK3 && scheme != SLES_SCHEME_ETD2)
    return fail("sles_step_drive_batch: scheme must be RK4, RK3 or ETD2");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_step_drive_batch: les_model must be 0, 1 or 2");
  if (members < 1 || members > SLES_BATCH_MAX_MEMBERS)
    return fail("sles_step_drive_batch: members must be in [1, " + std::to_string(SLES_BATCH_MAX_MEMBERS) + "]");
  if (prm->les_model && !sgs_coeff) return fail("sles_step_drive_batch: LES needs the per-member sgs_coeff");
  if (scheme != SLES_SCHEME_ETD2 && !u_hat0) return fail("sles_step_drive_batch: RK schemes need u_hat0");
  if (scheme == SLES_SCHEME_RK4 && !u_hat1) return fail("sles_step_drive_batch: RK4 needs u_hat1");
  if (scheme == SLES_SCHEME_ETD2 && !aux) return fail("sles_step_drive_batch: ETD2 needs aux");
  if (nsteps < 0) return fail("sles_step_drive_batch: nsteps must be >= 0");
  if (!sles_supported_n(p->n)) return fail("sles_step_drive_batch: unsupported grid size");
  if (nsteps == 0) return 0;
  const long sbytes = 3 * ncplx(p) * cplx_bytes(p);
  const long wstride = (long)batch_work_stride(p, prm->les_model);
  const long dstride = (long)(SLES_DRIVE_STATE_SIZE * sizeof(double));
  hipStream_t s = (hipStream_t)stream;
  DriveState ds{};
  if (drive_enter(p, scheme, prm->les_model, u_hat, u_hat0, u_hat1, aux, members, sbytes, state, dstride, work,
                  wstride, s, &ds))
    return -1;
  Batch b{};
  b.count = members;
  b.mb.state = ds.mstride;
  b.mb.work = wstride;
  b.mb.drive = dstride;
  b.mb.coeff = prm->les_model ? sgs_coeff : nullptr;
  DriveArgs drv{state, cfl_first};
  for (int i = 0; i < nsteps; ++i) {
    drv.xinv_in = i > 0;  // read only where the step fuses (RK4, fuse_x)
    drv.xinv_out = i + 1 < nsteps;
    const int rc = p->precision == SLES_FP64
                       ? step_f64(p, scheme, prm, ds.uh, ds.uh0, ds.uh1, ds.aux, ds.ps, ds.split, nullptr, nullptr,
                                  nullptr, work, s, nullptr, &drv, &b)
                       : step_f32(p, scheme, prm, ds.uh, ds.uh0, ds.uh1, ds.aux, ds.ps, ds.split, nullptr, nullptr,
                                  nullptr, work, s, nullptr, &drv, &b);
    if (rc) return rc;
  }
  return drive_exit(p, scheme, prm->les_model, u_hat, u_hat0, u_hat1, aux, members, sbytes, state, dstride, work,
                    wstride, s, ds);
}

int sles_rk4_stage_states(sles_plan p, const sles_step_params* prm, void* u_hat, void* u_hat0, void* u_hat1,
                          void* stages, void* work, void* stream) {
  if (!p || !prm || !u_hat || !u_hat0 || !u_hat1 || !stages || !work)
    return fail("sles_rk4_stage_states: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_rk4_stage_states: les_model must be 0, 1 or 2");
  if (prm->dt_device) return fail("sles_rk4_stage_states: needs a host dt (dt_device = 0)");
  if (!sles_supported_n(p->n)) return fail("sles_rk4_stage_states: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  const long ps = ncplx(p) / p->n;  // the caller's arrays: reference layout
  return p->precision == SLES_FP64
             ? step_f64(p, SLES_SCHEME_RK4, prm, u_hat, u_hat0, u_hat1, nullptr, ps, 0, nullptr, nullptr, nullptr,
                        work, s, stages)
             : step_f32(p, SLES_SCHEME_RK4, prm, u_hat, u_hat0, u_hat1, nullptr, ps, 0, nullptr, nullptr, nullptr,
                        work, s, stages);
}

int sles_profile_enable(int on) {
  g_prof = on != 0;
  return 0;
}

int sles_profile_read(double* ms, long* count, int nkinds) {
  if (!ms || !count || nkinds < 1) return fail("sles_profile_read: bad argument");
  int rc = 0;
  for (const ProfRec& r : g_recs) {
    float t = 0.f;
    if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) rc = -1;
    if (r.kind < nkinds) {
      ms[r.kind] += t;
      count[r.kind] += 1;
    }
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
  return rc ? fail("sles_profile_read: event query failed") : 0;
}

}  // extern "C"

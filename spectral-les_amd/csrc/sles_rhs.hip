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

using namespace sles;

extern "C" {

size_t sles_step_work_bytes(sles_plan p, int les_model) {
  if (!p) return 0;
  const size_t esz = p->precision == SLES_FP64 ? 16 : 8;
  const size_t nw = les_model ? 9 : 6;
  return nw * (size_t)ncplx(p) * esz + sizeof(double) * ((size_t)p->n * p->n + 64);
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
  return p->precision == SLES_FP64
             ? step_f64(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, u_out, curl_out, scalars, work, s, nullptr)
             : step_f32(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, u_out, curl_out, scalars, work, s, nullptr);
}

int sles_step_noalloc(sles_plan p, int scheme, const sles_step_params* prm, void* u_hat, void* u_hat0,
                      void* u_hat1, void* aux, void* u_out, void* curl_out, double* scalars, void* stream) {
  if (!p || !prm) return fail("sles_step_noalloc: null argument");
  const size_t need = sles_step_work_bytes(p, prm->les_model);
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
  const DriveArgs drv{state, cfl_first};
  for (int i = 0; i < nsteps; ++i) {
    // U and curl are outputs only (each step recomputes the physical fields
    // it needs from U_hat, and dt from the z pass's max u): the steps before
    // the last would write values the next step overwrites unread
    const bool last = i + 1 == nsteps;
    void* uo = last ? u_out : nullptr;
    void* co = last ? curl_out : nullptr;
    const int rc = p->precision == SLES_FP64
                       ? step_f64(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, uo, co, nullptr, work, s, nullptr, &drv)
                       : step_f32(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, uo, co, nullptr, work, s, nullptr, &drv);
    if (rc) return rc;
  }
  return 0;
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
  const long esz = p->precision == SLES_FP64 ? 16 : 8;
  Batch b{};
  b.count = members;
  b.mb.state = 3 * ncplx(p) * esz;
  b.mb.work = (long)batch_work_stride(p, prm->les_model);
  b.mb.drive = (long)(SLES_DRIVE_STATE_SIZE * sizeof(double));
  b.mb.coeff = prm->les_model ? sgs_coeff : nullptr;
  hipStream_t s = (hipStream_t)stream;
  const DriveArgs drv{state, cfl_first};
  for (int i = 0; i < nsteps; ++i) {
    const int rc =
        p->precision == SLES_FP64
            ? step_f64(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, nullptr, nullptr, nullptr, work, s, nullptr, &drv, &b)
            : step_f32(p, scheme, prm, u_hat, u_hat0, u_hat1, aux, nullptr, nullptr, nullptr, work, s, nullptr, &drv, &b);
    if (rc) return rc;
  }
  return 0;
}

int sles_rk4_stage_states(sles_plan p, const sles_step_params* prm, void* u_hat, void* u_hat0, void* u_hat1,
                          void* stages, void* work, void* stream) {
  if (!p || !prm || !u_hat || !u_hat0 || !u_hat1 || !stages || !work)
    return fail("sles_rk4_stage_states: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_rk4_stage_states: les_model must be 0, 1 or 2");
  if (prm->dt_device) return fail("sles_rk4_stage_states: needs a host dt (dt_device = 0)");
  if (!sles_supported_n(p->n)) return fail("sles_rk4_stage_states: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  return p->precision == SLES_FP64
             ? step_f64(p, SLES_SCHEME_RK4, prm, u_hat, u_hat0, u_hat1, nullptr, nullptr, nullptr, nullptr, work, s,
                        stages)
             : step_f32(p, SLES_SCHEME_RK4, prm, u_hat, u_hat0, u_hat1, nullptr, nullptr, nullptr, nullptr, work, s,
                        stages);
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

// Generic batched 3D real<->complex transforms of the reference's layout:
// physical (N, N, N) real, z fastest; spectral (N, N, N/2+1) complex [kx][ky][kz].
//
//   forward  = numpy.fft.rfftn(u, axes=(0,1,2))     : z r2c, then y, then x
//   backward = numpy.fft.irfftn(U, s=(N,)*3, ...)    : x, then y, then z c2r
//
// (the reference's fft.forward / fft.backward, initialize_domain.py:59-65).
// Each pass reads and writes one full array; the x/y passes run along a
// strided axis with a tile of TK consecutive columns per block so every
// global access is a run of TK contiguous complex values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sles.h"
#include "sles_internal.h"

namespace sles {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}
int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// ---------------------------------------------------------------------------
// Contiguous-axis passes (two real rows per complex line).
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(ZShape<N>::THREADS)
    k_z_r2c(const T* __restrict__ in, cpx<T>* __restrict__ out, long nrows, T scale,
            const cpx<T>* __restrict__ tw) {
  using S = ZShape<N>;
  constexpr int P = S::P, NT = S::NT, NH = S::NH;
  __shared__ cpx<T> lds[S::LPB * LdsLine<T, N>::STRIDE];
  const LdsLine<T, N> map{(int)(threadIdx.x / S::NT) * LdsLine<T, N>::STRIDE};
  const int line = threadIdx.x / NT, t = threadIdx.x % NT;
  const TwRegs<T, LinePlan<N>> twr(tw, t);  // issued together with the data loads
  const long ra = 2 * ((long)blockIdx.x * S::LPB + line), rb = ra + 1;
  const bool va = ra < nrows, vb = rb < nrows;
  const long rla = va ? ra : 0, rlb = vb ? rb : 0;  // clamped rows (in bounds)
  cpx<T> v[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int z = t + m * NT;
    const T xa = in[rla * N + z], xb = in[rlb * N + z];
    v[m].x = va ? xa : T(0);
    v[m].y = vb ? xb : T(0);
  }
  LineFFTP<T, LinePlan<N>, -1, WaveLocal<LinePlan<N>>::value>::run(v, lds, map, t, twr);
  // Split Z = FFT(a + i b):  A_k = (Z_k + conj Z_{N-k})/2,  B_k = (Z_k - conj Z_{N-k})/(2i),
  // through the unpermuted layout of the line (mirrored reads)
  const LdsPlain pmap = map.plain();
#pragma unroll
  for (int m = 0; m < P; ++m) lds[pmap(t + m * NT)] = v[m];
  line_sync<WaveLocal<LinePlan<N>>::value>();
#pragma unroll
  for (int m = 0; m <= P / 2; ++m) {
    const int k = t + m * NT;
    if (k <= N / 2) {
      const cpx<T> zk = v[m];
      const cpx<T> zn = conj(lds[pmap((N - k) % N)]);
      const cpx<T> a = T(0.5) * (zk + zn);
      const cpx<T> d = zk - zn;
      const cpx<T> b = {T(0.5) * d.y, T(-0.5) * d.x};
      if (va) out[ra * NH + k] = scale * a;
      if (vb) out[rb * NH + k] = scale * b;
    }
  }
}

template <typename T, int N>
__global__ void __launch_bounds__(ZShape<N>::THREADS)
    k_z_c2r(const cpx<T>* __restrict__ in, long plane, int split, T* __restrict__ out, long nrows, T scale,
            const cpx<T>* __restrict__ tw) {
  using S = ZShape<N>;
  constexpr int P = S::P, NT = S::NT, NH = S::NH;
  __shared__ cpx<T> lds[S::LPB * LdsLine<T, N>::STRIDE];
  const LdsLine<T, N> map{(int)(threadIdx.x / S::NT) * LdsLine<T, N>::STRIDE};
  const int line = threadIdx.x / NT, t = threadIdx.x % NT;
  const TwRegs<T, LinePlan<N>> twr(tw, t);  // issued together with the data loads
  const long ra = 2 * ((long)blockIdx.x * S::LPB + line), rb = ra + 1;
  const bool va = ra < nrows, vb = rb < nrows;
  const long rla = va ? ra : 0, rlb = vb ? rb : 0;  // clamped rows (in bounds)
  constexpr bool WV = WaveLocal<LinePlan<N>>::value;
  // Each half-spectrum element is read from memory once (bins k <= N/2: slots
  // m < P/2, plus k = N/2 in thread 0); the Hermitian completion k > N/2 takes
  // conj of bin N - k from the line's LDS row (unpermuted layout).
  cpx<T> ha[P / 2 + 1], hb[P / 2 + 1];
  // bin k of row r: plane r / N; row r % N at (r % N) (N/2+1) + k, or split:
  // at (r % N) N/2 + k for k < N/2 and N N/2 + r % N for k = N/2
  auto bin = [&](long r, int k) {
    const long x = r / N, y = r % N;
    if (!split) return x * plane + y * NH + k;
    return x * plane + (k < N / 2 ? y * (N / 2) + k : (long)N * (N / 2) + y);
  };
#pragma unroll
  for (int m = 0; m <= P / 2; ++m) {
    const int k = t + m * NT;
    if (k <= N / 2) {
      ha[m] = in[bin(rla, k)];
      hb[m] = in[bin(rlb, k)];
      if (!va) ha[m] = {T(0), T(0)};
      if (!vb) hb[m] = {T(0), T(0)};
    }
  }
  const LdsPlain pmap = map.plain();
  cpx<T> v[P];
  auto complete = [&](const cpx<T>(&h)[P / 2 + 1], bool second) {
#pragma unroll
    for (int m = 0; m <= P / 2; ++m)
      if (t + m * NT <= N / 2) lds[pmap(t + m * NT)] = h[m];
    line_sync<WV>();
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int k = t + m * NT;
      cpx<T> x = lds[pmap(k <= N / 2 ? k : N - k)];
      if (k > N / 2)
        x = conj(x);
      else if (k == 0 || k == N / 2)  // c2r ignores the imaginary part of these bins
        x.y = T(0);
      v[m] = second ? cpx<T>{v[m].x - x.y, v[m].y + x.x} : x;  // a + i b
    }
    line_sync<WV>();
  };
  complete(ha, false);
  complete(hb, true);
  LineFFTP<T, LinePlan<N>, +1, WV>::run(v, lds, map, t, twr);
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int z = t + m * NT;
    if (va) out[ra * N + z] = scale * v[m].x;
    if (vb) out[rb * N + z] = scale * v[m].y;
  }
}

// ---------------------------------------------------------------------------
// Strided complex pass: batch of fields, each viewed as (outer, N, M) with the
// transform along the middle axis.  Column g in [0, outer*M) -> (g / M, g % M).
// One block per tile of TK columns runs every field of the batch, the next
// field's loads issued before the current field's transform.  S: tile shape,
// PL: line plan, TWR: twiddles in registers; else twiddles from the table and
// one field per block (grid.y = fields).
// ---------------------------------------------------------------------------
template <typename T, int N, int DIR, class S = WideShape<T, N>, class PL = LinePlan<N>, bool TWR = true>
__global__ void __launch_bounds__((S::THREADS))
    k_strided(const cpx<T>* in, cpx<T>* out, long outer, long M, long fstride, int nfields, T scale,
              const cpx<T>* __restrict__ tw, int xcd) {
  constexpr int P = PL::P, NT = PL::NT, TK = S::TK;
  static_assert(S::P == P && S::THREADS == TK * NT, "tile shape and line plan disagree");
  __shared__ cpx<T> lds[N * TK];
  const int c = threadIdx.x % TK, t = threadIdx.x / TK;
  const auto twr = TwPick<T, PL, TWR>::make(tw, t);  // issued together with the data loads
  const long g = tile_id(xcd) * TK + c;
  const bool valid = g < outer * M;
  const long o = valid ? g / M : 0, col = valid ? g % M : 0;
  const long base = o * (long)N * M + col;
  cpx<T> v[P], nx[P];
  if constexpr (!TWR) {
    // Table-twiddle (wide) tiles: one field per block (blockIdx.y), element
    // offsets within the field in 32 bits (a field is below 4 GiB: N <= 512)
    // and a uniform field base pointer, to stay inside 128 VGPRs.
    const unsigned b32 = (unsigned)base, m32 = (unsigned)M;
    const long fo = (long)blockIdx.y * fstride;
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = ld32(in + fo, b32 + (unsigned)(t + m * NT) * m32);
    LineFFTP<T, PL, DIR>::S::run(v, lds, LdsColumn<TK>{c}, t, TwTable32<T>{tw});
    if (valid) {
#pragma unroll
      for (int m = 0; m < P; ++m) st32(out + fo, b32 + (unsigned)(t + m * NT) * m32, scale * v[m]);
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < P; ++m) nx[m] = in[base + (long)(t + m * NT) * M];  // in bounds (clamped); unused if !valid
#pragma unroll 1
  for (int f = 0; f < nfields; ++f) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = nx[m];
    if (f + 1 < nfields) {
#pragma unroll
      for (int m = 0; m < P; ++m) nx[m] = in[(f + 1) * fstride + base + (long)(t + m * NT) * M];
    }
    LineFFTP<T, PL, DIR>::run(v, lds, LdsColumn<TK>{c}, t, twr);
    if (valid) {
#pragma unroll
      for (int m = 0; m < P; ++m) out[f * fstride + base + (long)(t + m * NT) * M] = scale * v[m];
    }
  }
}

// Wide tiles for N = 256, 512: twice the columns of WideShape (LDS tile up to
// 128 KiB, 1024 threads), so every row of a tile is a run of 256 B; twiddles
// come from the table to stay inside the 128 VGPRs of a 1024-thread block.
// FP32 at N = 512 holds 16 points per thread (radix-8 stages, two butterflies
// per thread per stage) to keep 32 columns within 1024 threads.
template <typename T, int N>
struct WidePlan {
  using type = LinePlan<N>;
};
template <>
struct WidePlan<float, 512> {
  using type = Plan<512, 16, 8, 8, 8>;
};
template <typename T, int N>
struct WideTile {
  using PL = typename WidePlan<T, N>::type;
  static constexpr int P = PL::P;
  static constexpr int NT = PL::NT;
  static constexpr int TKL = 131072 / (N * (int)sizeof(cpx<T>));
  static constexpr int TKT = 1024 / NT;
  static constexpr int TK = TKL < TKT ? TKL : TKT;
  static constexpr int THREADS = TK * NT;
};
template <int N>
struct HasWide {
  static constexpr bool value = N == 256 || N == 512;
};

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <typename T, int N>
static int launch_z_r2c(const sles_plan_s* p, const void* in, void* out, long nrows, double scale,
                        hipStream_t s) {
  using S = ZShape<N>;
  const long pairs = (nrows + 1) / 2;
  const long blocks = (pairs + S::LPB - 1) / S::LPB;
  hipLaunchKernelGGL((k_z_r2c<T, N>), dim3((unsigned)blocks), dim3(S::THREADS), 0, s, (const T*)in,
                     (cpx<T>*)out, nrows, (T)scale, (const cpx<T>*)p->tw);
  return check_launch("z_r2c");
}

template <typename T, int N>
static int launch_z_c2r(const sles_plan_s* p, const void* in, long plane, int split, void* out, long nrows,
                        double scale, hipStream_t s) {
  using S = ZShape<N>;
  const long pairs = (nrows + 1) / 2;
  const long blocks = (pairs + S::LPB - 1) / S::LPB;
  hipLaunchKernelGGL((k_z_c2r<T, N>), dim3((unsigned)blocks), dim3(S::THREADS), 0, s,
                     (const cpx<T>*)in, plane, split, (T*)out, nrows, (T)scale, (const cpx<T>*)p->tw);
  return check_launch("z_c2r");
}

template <typename T, int N, class S, class PL, bool TWR>
static void launch_strided_shape(const sles_plan_s* p, const void* in, void* out, long outer, long M,
                                 long fstride, int nfields, int dir, double scale, hipStream_t s) {
  const long blocks = (outer * M + S::TK - 1) / S::TK;
  // table-twiddle tiles run one field per block row (grid.y)
  const dim3 grid((unsigned)blocks, TWR ? 1u : (unsigned)nfields);
  if (dir < 0)
    hipLaunchKernelGGL((k_strided<T, N, -1, S, PL, TWR>), grid, dim3(S::THREADS), 0, s,
                       (const cpx<T>*)in, (cpx<T>*)out, outer, M, fstride, nfields, (T)scale,
                       (const cpx<T>*)p->tw, p->tune_xcd);
  else
    hipLaunchKernelGGL((k_strided<T, N, +1, S, PL, TWR>), grid, dim3(S::THREADS), 0, s,
                       (const cpx<T>*)in, (cpx<T>*)out, outer, M, fstride, nfields, (T)scale,
                       (const cpx<T>*)p->tw, p->tune_xcd);
}

// wide: use WideTile (sizes that have it)
template <typename T, int N>
static int launch_strided(const sles_plan_s* p, const void* in, void* out, long outer, long M,
                          long fstride, int nfields, int dir, double scale, bool wide, hipStream_t s) {
  if constexpr (HasWide<N>::value) {
    if (wide && nfields <= 65535) {
      using W = WideTile<T, N>;
      launch_strided_shape<T, N, W, typename W::PL, false>(p, in, out, outer, M, fstride, nfields, dir, scale, s);
      return check_launch("strided (wide)");
    }
  }
  launch_strided_shape<T, N, WideShape<T, N>, LinePlan<N>, true>(p, in, out, outer, M, fstride, nfields, dir,
                                                                 scale, s);
  return check_launch("strided");
}

template <typename T>
static int fft_r2c_t(const sles_plan_s* p, const void* in, void* out, int batch, double scale,
                     hipStream_t s) {
  const int n = p->n;
  const long C = ncplx(p);
  const long rows = (long)batch * n * n;
  const bool wx = (p->tune_fft_wide & 1) != 0, wy = (p->tune_fft_wide & 2) != 0;
  int rc = 0;
#define SLES_CASE(NN)                                                                     \
  case NN:                                                                                \
    rc = launch_z_r2c<T, NN>(p, in, out, rows, 1.0, s);                                   \
    if (!rc) rc = launch_strided<T, NN>(p, out, out, (long)n, (long)p->nh, C, batch, -1, 1.0, wy, s); \
    if (!rc) rc = launch_strided<T, NN>(p, out, out, 1, (long)n * p->nh, C, batch, -1, scale, wx, s); \
    break;
  switch (n) {
    SLES_FOR_EACH_N(SLES_CASE)
    default:
      return fail("unsupported grid size");
  }
#undef SLES_CASE
  return rc;
}

template <typename T>
static int fft_c2r_t(const sles_plan_s* p, const void* in, void* out, void* work, int batch,
                     double scale, hipStream_t s) {
  const int n = p->n;
  const long C = ncplx(p);
  const long rows = (long)batch * n * n;
  const bool wx = (p->tune_fft_wide & 1) != 0, wy = (p->tune_fft_wide & 2) != 0;
  int rc = 0;
#define SLES_CASE(NN)                                                                      \
  case NN:                                                                                 \
    rc = launch_strided<T, NN>(p, in, work, 1, (long)n * p->nh, C, batch, +1, 1.0, wx, s);  \
    if (!rc) rc = launch_strided<T, NN>(p, work, work, (long)n, (long)p->nh, C, batch, +1, 1.0, wy, s); \
    if (!rc) rc = launch_z_c2r<T, NN>(p, work, (long)n * p->nh, 0, out, rows, scale, s);   \
    break;
  switch (n) {
    SLES_FOR_EACH_N(SLES_CASE)
    default:
      return fail("unsupported grid size");
  }
#undef SLES_CASE
  return rc;
}

// Internal entry points shared with the fused time-step code.
int z_c2r_rows(const sles_plan_s* p, const void* in, long plane, int split, void* out, long nrows, double scale,
               hipStream_t s) {
  int rc = 0;
#define SLES_CASE(NN)                                                                         \
  case NN:                                                                                    \
    rc = p->precision == SLES_FP64 ? launch_z_c2r<double, NN>(p, in, plane, split, out, nrows, scale, s) \
                                   : launch_z_c2r<float, NN>(p, in, plane, split, out, nrows, scale, s);  \
    break;
  switch (p->n) {
    SLES_FOR_EACH_N(SLES_CASE)
    default:
      return fail("unsupported grid size");
  }
#undef SLES_CASE
  return rc;
}

}  // namespace sles

using namespace sles;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* sles_last_error(void) { return g_err.c_str(); }

int sles_version(void) { return SLES_ABI_VERSION; }

int sles_plan_create(sles_plan* out, int n, int precision) {
  if (!out) return fail("sles_plan_create: null output pointer");
  *out = nullptr;
  if (!sles_supported_n(n)) return fail("sles_plan_create: unsupported grid size " + std::to_string(n));
  if (precision != SLES_FP64 && precision != SLES_FP32)
    return fail("sles_plan_create: precision must be SLES_FP64 or SLES_FP32");
  sles_plan_s* p = new sles_plan_s();
  p->n = n;
  p->nh = n / 2 + 1;
  p->precision = precision;
  // wide tiles of the generic transforms where they measured faster (MI355X,
  // batch of 3 fields): both passes at 512, the x pass (FP64) or the y pass
  // (FP32) at 256
  p->tune_fft_wide = n == 512 ? 3 : n == 256 ? (precision == SLES_FP64 ? 1 : 2) : 0;
  // fused-step workspace planes padded by 5 lines (640 B) from N = 128 on,
  // where the strided x passes stream faster with it (256^3 FP64 access
  // pattern of the x-forward pass 3.8 -> 4.9 TB/s); the small grids are
  // launch bound and keep the compact layout
  p->tune_pad = n >= 128 ? 5 : 0;
  (void)hipGetDevice(&p->device);
  step_scratch_init(p->device);
  // Twiddles exp(-2 pi i e/n) evaluated in long double and rounded once.
  const size_t esz = precision == SLES_FP64 ? 16 : 8;
  std::vector<unsigned char> host(esz * n);
  for (int e = 0; e < n; ++e) {
    const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
    const long double c = cosl(ang), sn = sinl(ang);
    if (precision == SLES_FP64) {
      double v[2] = {(double)c, (double)sn};
      std::memcpy(&host[esz * e], v, esz);
    } else {
      float v[2] = {(float)c, (float)sn};
      std::memcpy(&host[esz * e], v, esz);
    }
  }
  if (hipMalloc(&p->tw, esz * n) != hipSuccess) {
    delete p;
    return fail("sles_plan_create: hipMalloc failed for twiddle table");
  }
  if (hipMemcpy(p->tw, host.data(), esz * n, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(p->tw);
    delete p;
    return fail("sles_plan_create: twiddle upload failed");
  }
  *out = p;
  return 0;
}

int sles_plan_destroy(sles_plan p) {
  if (!p) return 0;
  if (p->tw) (void)hipFree(p->tw);
  delete p;
  return 0;
}

int sles_plan_info(sles_plan p, int* n, int* nh, int* precision) {
  if (!p) return fail("sles_plan_info: null plan");
  if (n) *n = p->n;
  if (nh) *nh = p->nh;
  if (precision) *precision = p->precision;
  return 0;
}

int sles_plan_set_tuning(sles_plan p, int key, int value) {
  if (!p) return fail("sles_plan_set_tuning: null plan");
  switch (key) {
    case SLES_TUNE_STRIDED_RADIX4: p->tune_strided = value != 0; return 0;
    case SLES_TUNE_XFWD_RADIX4: p->tune_xfwd = value != 0; return 0;
    case SLES_TUNE_Z_RADIX4: p->tune_z = value != 0; return 0;
    case SLES_TUNE_XCD: p->tune_xcd = value != 0; return 0;
    case SLES_TUNE_FUSE_X: p->tune_fuse_x = value != 0; return 0;
    case SLES_TUNE_FFT_WIDE:
      if (value < 0 || value > 3) return fail("sles_plan_set_tuning: SLES_TUNE_FFT_WIDE takes 0..3");
      p->tune_fft_wide = value;
      return 0;
    case SLES_TUNE_PAD:
      if (value < 0 || value > SLES_TUNE_PAD_MAX)
        return fail("sles_plan_set_tuning: SLES_TUNE_PAD takes 0.." + std::to_string(SLES_TUNE_PAD_MAX));
      p->tune_pad = value;
      return 0;
    case SLES_TUNE_NARROW: p->tune_narrow = value != 0; return 0;
    default: return fail("sles_plan_set_tuning: unknown key");
  }
}

int sles_plan_get_tuning(sles_plan p, int key, int* value) {
  if (!p || !value) return fail("sles_plan_get_tuning: null argument");
  switch (key) {
    case SLES_TUNE_STRIDED_RADIX4: *value = p->tune_strided; return 0;
    case SLES_TUNE_XFWD_RADIX4: *value = p->tune_xfwd; return 0;
    case SLES_TUNE_Z_RADIX4: *value = p->tune_z; return 0;
    case SLES_TUNE_XCD: *value = p->tune_xcd; return 0;
    case SLES_TUNE_FFT_WIDE: *value = p->tune_fft_wide; return 0;
    case SLES_TUNE_FUSE_X: *value = p->tune_fuse_x; return 0;
    case SLES_TUNE_PAD: *value = p->tune_pad; return 0;
    case SLES_TUNE_NARROW: *value = p->tune_narrow; return 0;
    default: return fail("sles_plan_get_tuning: unknown key");
  }
}

int sles_fft_r2c(sles_plan p, const void* in, void* out, int batch, double scale, void* stream) {
  if (!p || !in || !out) return fail("sles_fft_r2c: null argument");
  if (batch < 1) return fail("sles_fft_r2c: batch must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  return p->precision == SLES_FP64 ? fft_r2c_t<double>(p, in, out, batch, scale, s)
                                   : fft_r2c_t<float>(p, in, out, batch, scale, s);
}

int sles_fft_c2r(sles_plan p, const void* in, void* out, void* work, int batch, double scale,
                 void* stream) {
  if (!p || !in || !out || !work) return fail("sles_fft_c2r: null argument");
  if (batch < 1) return fail("sles_fft_c2r: batch must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  return p->precision == SLES_FP64 ? fft_c2r_t<double>(p, in, out, work, batch, scale, s)
                                   : fft_c2r_t<float>(p, in, out, work, batch, scale, s);
}

}  // extern "C"

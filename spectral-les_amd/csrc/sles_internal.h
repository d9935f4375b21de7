// Internal declarations shared by the libsles translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "sles_fft.h"

struct sles_plan_s {
  int n;          // grid points per axis (cubic grid)
  int nh;         // n/2 + 1 spectral bins on the last axis
  int precision;  // 0 = FP64, 1 = FP32
  int device;
  void* tw;       // device table tw[e] = exp(-2 pi i e/n), e in [0, n)
  // kernel variants (sles_plan_set_tuning): 0 = radix-8 line plan, 1 = radix-4
  int tune_strided = 0;  // x/y inverse and y forward passes
  int tune_xfwd = 0;     // x forward + stage update
  int tune_z = 0;        // fused z kernel
  int tune_xcd = 1;      // XCD-aware tile order in the strided passes (default on)
  int tune_fuse_x = 1;   // RK4 x-forward runs the next stage's x-inverse where compiled
  int tune_fft_wide = 0; // generic transforms: wide tiles in the x (bit 0) / y (bit 1) pass
                         // (default set per size in sles_plan_create)
  int tune_pad = 0;      // fused-step workspace: kx planes padded by this many 128-byte lines
                         // (default set per size in sles_plan_create)
  int tune_narrow = 0;   // fused step, N = 64..128: narrow small-grid tiles in the strided passes
};

namespace sles {

void set_error(const std::string& msg);
int fail(const std::string& msg);
int check_launch(const char* what);

inline long nreal(const sles_plan_s* p) { return (long)p->n * p->n * p->n; }
inline long ncplx(const sles_plan_s* p) { return (long)p->n * p->n * p->nh; }
inline long cplx_bytes(const sles_plan_s* p) { return p->precision == 0 ? 16 : 8; }  // one spectral value

// Fused-step workspace layout: kx planes of N (N/2+1) spectral values start
// every plane_w elements (padded by tune_pad 128-byte lines: unpadded planes
// are a multiple of 4 KiB apart at N = 256 and 512, and the strided x passes,
// which read 128-256-byte runs from every plane, then stream ~25 % slower),
// fields every cplx_w = N plane_w elements.
inline long plane_w(const sles_plan_s* p) {
  return (long)p->n * p->nh + (long)p->tune_pad * 128 / cplx_bytes(p);
}
inline long cplx_w(const sles_plan_s* p) { return (long)p->n * plane_w(p); }

// Wavenumbers of the reference layout (initialize_domain.py:76-79,99-107):
// axes 0 and 1 use fftfreq(N, 1/N) (index N/2 -> -N/2); axis 2 holds 0..N/2
// with the last entry +N/2.
__device__ __forceinline__ int kfreq(int i, int n) { return i < n / 2 ? i : i - n; }

// Spectral element e of a (N, N, N/2+1) array -> (kx, ky, kz).
struct KIndex {
  int kx, ky, kz;
};
__device__ __forceinline__ KIndex kindex(long e, int n, int nh) {
  const int iz = (int)(e % nh);
  const long r = e / nh;
  const int iy = (int)(r % n);
  const int ix = (int)(r / n);
  return {kfreq(ix, n), kfreq(iy, n), iz};
}

// dealias = prod_i |K_i| < 2/3 (N//2 + 1)   (initialize_domain.py:79,109-114)
__device__ __forceinline__ bool dealias_keep(KIndex k, int n) {
  const double kmax = 2.0 / 3.0 * (double)(n / 2 + 1);
  return (double)abs(k.kx) < kmax && (double)abs(k.ky) < kmax && (double)abs(k.kz) < kmax;
}

// Tile of this block in a strided pass.  With xcd != 0 (and a grid that is a
// multiple of 8) consecutive tiles go to the same XCD: hardware dispatches
// block b to XCD b mod 8, so tile (b mod 8) * (grid/8) + b/8 keeps
// neighbouring columns -- which share cache lines on the y axis -- inside one
// L2.
__device__ __forceinline__ long tile_id(int xcd) {
  const long b = blockIdx.x;
  if (xcd && (gridDim.x & 7) == 0) return (b & 7) * (long)(gridDim.x >> 3) + (b >> 3);
  return b;
}

// Shapes of the generic passes.
template <int N>
struct ZShape {
  static constexpr int P = LinePlan<N>::P;
  static constexpr int NT = N / P;
  static constexpr int LPB = (256 / NT) > 0 ? (256 / NT) : 1;  // lines per block
  static constexpr int THREADS = LPB * NT;
  static constexpr int NH = N / 2 + 1;
};

constexpr long pow2_floor(long v) { return v < 2 ? 1 : 2 * pow2_floor(v / 2); }

// Line plan PL whose strided passes take the narrow small-grid tiles (plan
// knob SLES_TUNE_NARROW): four tiles per CU instead of one.  A tile's width
// changes which block runs a column, not the arithmetic on it.
template <class PL>
struct Narrow : PL {};
template <class PL>
struct IsNarrow {
  static constexpr bool value = false;
};
template <class PL>
struct IsNarrow<Narrow<PL>> {
  static constexpr bool value = true;
};

// Strided passes: TK columns per block, LDS tile N*TK complex <= 64 KiB.
// Blocks of plans with >= 6 points per thread are capped at 512 threads (8
// waves: up to 256 VGPRs per wave without spilling); radix-4 plans (4 points
// per thread, about half the registers) may use 1024.  MAXT_ > 0 overrides
// the cap for one kernel.  Small grids have few columns (N (N/2+1): 2112 at
// 64^3), so tiles of the LDS width would leave most CUs idle (33 blocks at
// 64^3 FP64); there a tile spans at most N (N/2+1) / 256 columns (one tile
// per CU; Narrow plans: / 1024, four per CU), but at least one wavefront's
// worth (64 / NT).
template <typename T, int N, class PL = LinePlan<N>, int MAXT_ = 0>
struct StridedShape {
  static constexpr int P = PL::P;
  static constexpr int NT = N / P;
  static constexpr int MAXT = MAXT_ > 0 ? MAXT_ : (P >= 6 ? 512 : 1024);
  static constexpr int TK0 = 65536 / (N * (int)sizeof(cpx<T>));
  static constexpr int FILL0 = (int)pow2_floor((long)N * (N / 2 + 1) / (IsNarrow<PL>::value ? 1024 : 256));
  static constexpr int FILL = FILL0 > 64 / NT ? FILL0 : (64 / NT > 0 ? 64 / NT : 1);
  static constexpr int TKL = TK0 > 64 ? 64 : TK0;
  static constexpr int TK1 = TKL < FILL ? TKL : FILL;
  static constexpr int TK = (TK1 * NT > MAXT) ? (MAXT / NT > 0 ? MAXT / NT : 1) : TK1;
  static constexpr int THREADS = TK * NT;
};

// Strided passes of FP32 lines with >= 64 threads per line (radix-8 plans at
// N >= 512): 1024-thread blocks, so that a tile still spans 16 columns
// (128 B runs per row) -- their register use stays far below the 128-VGPR
// limit of such blocks.  Other cases keep the default cap.
template <typename T, int N, class PL = LinePlan<N>>
using WideShape = StridedShape<T, N, PL, (sizeof(T) == 4 && PL::NT >= 64) ? 1024 : 0>;

// Takes the address of the no-allocation step's device scratch on `device`
// (sles_rhs.hip); sles_plan_create calls it with that device current.
void step_scratch_init(int device);

// c2r along z of nrows half-spectrum rows (row r: plane r / N, row r % N of
// planes `plane` elements apart, arranged as the reference's (split = 0) or
// as the fused step's workspace (split = 1: rows of N/2, then the Nyquist
// column)) into nrows real rows of N.
int z_c2r_rows(const sles_plan_s* p, const void* in, long plane, int split, void* out, long nrows, double scale,
               hipStream_t s);

}  // namespace sles

// Sizes with a compiled line plan.
#define SLES_FOR_EACH_N(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

inline bool sles_supported_n(int n) {
#define SLES_CASE(NN) if (n == NN) return true;
  SLES_FOR_EACH_N(SLES_CASE)
#undef SLES_CASE
  return false;
}

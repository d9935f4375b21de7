// Fused RHS + time-step kernels.
//
// One right-hand-side evaluation of RK4_method.computeRHS
// (src/RK4_method.py:38-71, with Smagorinsky_model / Vreman_model,
// src/smagorinsky.py:8-47, src/vreman.py:8-123, and math_formula.py) plus the
// stage update runs as five transform passes over a workspace W of up to 9
// spectral fields.  Wavenumbers come from indices; the physical fields never
// touch HBM.
//
//   LES (Smagorinsky / Vreman)              no LES
//   1 x-inverse  U_hat(3) -> W0-2 = u, W3-5 = i Kx u       W0-2 = u, W3-4 = i Kx v, i Kx w
//   2 y-inverse  W0-5 -> + W6-8 = i Ky u (9 fields)        5 fields -> W0-2 = u, W3-5 = i K x u
//   3 z kernel   9 half-spectra -> c2r of u, du/dx, du/dy, du/dz (12 fields)
//                -> omega, u x omega, nu_R, tau = nu_R S, max(u)
//                -> r2c -> W0-2 = F(u x omega)_i + 2i Kz tau_iz,
//                          W3-7 = tau_xx, tau_xy, tau_xz, tau_yy, tau_yz
//                                                          c2r u, omega -> u x omega -> W0-2
//   4 y-forward  W0-2 += 2i Ky tau_iy (6 fields out)       3 fields in place
//   5 x-forward  S_i = dealias (F_x(W_i) + 2i Kx F_x(tau_xi)); projection,
//                forcing, -nu K^2 U_hat, and the stage update of the scheme.
//
// Linear algebra: dU = dealias F(u x omega) + T - P_hat K + f U_hat - nu K^2 U_hat
// with T = 2i dealias sum_j K_j F(nu_R S_ij) and P_hat = sum_i (dU + T)_i K_i/K^2
// is the value the reference computes; the sums are regrouped (K_j commutes
// with the transforms along the other axes).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "sles.h"
#include "sles_internal.h"

namespace sles {

enum { EPI_RK4 = 0, EPI_RK3 = 1, EPI_ETD2 = 2, EPI_RHS = 3, EPI_NONLIN = 4 };

// Slots of the sles_step_drive state (the SLES_DRIVE_STATE_* layout of sles.h).
enum {
  DRV_UMAX = SLES_DRIVE_STATE_UMAX,
  DRV_DT = SLES_DRIVE_STATE_DT,
  DRV_T = SLES_DRIVE_STATE_T,
  DRV_STATUS = SLES_DRIVE_STATE_STATUS,
  DRV_STEPS = SLES_DRIVE_STATE_STEPS,
  DRV_DT_APPLY = SLES_DRIVE_STATE_DT_APPLY,
  DRV_T_PAUSE = SLES_DRIVE_STATE_T_PAUSE,
  DRV_T_FINAL = SLES_DRIVE_STATE_T_FINAL
};

// Element index inside one spectral field and the accessors of the strided
// passes.
using eidx = long;
template <typename T>
__device__ __forceinline__ T ldx(const T* b, eidx i) { return b[i]; }
template <typename T>
__device__ __forceinline__ void stx(T* b, eidx i, T v) { b[i] = v; }

struct StepArgs {
  double dt;             // host step size (when dt_dev == nullptr)
  const double* dt_dev;  // on-device step size, or nullptr
  double nu, forcing;
  double ca, cb;  // RK4: a[s], b[s]
  int scheme, stage;
  void* stages;  // RK4: if set, stage s < 3 also writes its new state Y_{s+1} to stages + 3 s C
};

// sles_step_drive: the driver-loop state (device doubles, DRV_* layout) and
// the CFL of the first step.  Between two RK4 steps of one driver call, the
// last stage's x-forward can run the next step's first x-inverse (the plan's
// fuse_x knob): xinv_out asks a step to do so, xinv_in tells a step that its
// workspace already holds that x-inverse (the previous step was asked).
struct DriveArgs {
  double* state;
  double cfl_first;
  bool xinv_in = false, xinv_out = false;
};

// Batched launches (sles_step_drive_batch): blockIdx.y is the member, whose
// U_hat (and U_hat0, U_hat1, aux), workspace (W and the CFL partials behind
// it) and driver state lie at fixed byte strides from the first member's.
// Single-member launches pass zero strides and a one-row grid.
struct Members {
  long state;           // bytes between members' U_hat (also U_hat0, U_hat1, aux)
  long work;            // bytes between members' workspaces
  long drive;           // bytes between members' driver states
  const double* coeff;  // per-member SGS coefficient (device), or nullptr: the launch's own
};

// The members of a batched step: count grid rows at the strides of mb.
struct Batch {
  int count;
  Members mb;
};

template <class P>
__device__ __forceinline__ P* member_ptr(P* p, long stride) {
  using B = typename std::conditional<std::is_const<P>::value, const char, char>::type;
  return reinterpret_cast<P*>(reinterpret_cast<B*>(p) + (long)blockIdx.y * stride);
}

// ---------------------------------------------------------------------------
// Tile geometry of the strided passes (N * (N/2+1) columns, TK per block).
// ---------------------------------------------------------------------------
// Tiles are numbered by tile_id (sles_internal.h), XCD-aware when xcd != 0.

template <int N, int TK>
struct Cols {
  static constexpr int NH = N / 2 + 1;
  static constexpr long M = (long)N * NH;
  int c;    // column of the tile
  int t;    // the thread's slot in that column's line (points t + m NT)
  long g0;  // the tile's first column (uniform over the block)
  long g;
  bool valid;
  // Lanes run across the tile's columns (c = threadIdx.x mod TK), so that a
  // row of the tile is one contiguous run -- or, with `along` (uniform over
  // the block), along each column's line (t = threadIdx.x mod nt), for
  // tiles whose lines are contiguous in memory.
  __device__ __forceinline__ explicit Cols(int xcd, int nt = 1, bool along = false) {
    if (along) {
      c = threadIdx.x / nt;
      t = threadIdx.x % nt;
    } else {
      c = threadIdx.x % TK;
      t = threadIdx.x / TK;
    }
    g0 = tile_id(xcd) * TK;
    g = g0 + c;
    valid = g < M;
    if (!valid) g = 0;
  }
};


// Arrangement of one kx plane of the workspace (and of the drivers' copies
// of the RK state): the rows kz < N/2 of every ky first (ky * N/2 + kz), then
// the N values of the Nyquist column kz = N/2 (N * N/2 + ky).  Every row the
// y passes and the z pass touch then starts on a multiple of N/2 values
// (aligned 128-byte runs) instead of every N/2 + 1, and the x passes' tiles of
// consecutive plane columns are still contiguous runs.  The reference layout
// (ky * (N/2+1) + kz) stays the caller's.
template <int N>
struct Split {
  static constexpr int NR = N / 2, NH = N / 2 + 1;
  static constexpr long MAIN = (long)N * NR;  // the rows; the Nyquist column follows
  // (ky index, kz) of plane column j
  __device__ __forceinline__ static void at(long j, int& iy, int& kz) {
    if (j < MAIN) {
      iy = (int)(j / NR);
      kz = (int)(j % NR);
    } else {
      iy = (int)(j - MAIN);
      kz = NR;
    }
  }
  // plane column of (ky index, kz)
  __device__ __forceinline__ static long col(int iy, int kz) { return kz < NR ? (long)iy * NR + kz : MAIN + iy; }
};

// Column g of an x-pass tile.  With the RK state in the workspace
// arrangement (a driver's copies, ssplit != 0) the tiles run over the split
// plane's columns, so the workspace and the state accesses are both
// contiguous runs; with the state in the reference layout (the caller's
// arrays) they run over the reference plane's columns, which keeps those
// runs contiguous, and the workspace column follows.
template <int N>
struct XCol {
  long jw, js;  // plane column in the workspace, in the state arrays
  int iy, kz;   // ky index and kz of the column
  __device__ __forceinline__ XCol(long g, int ssplit) {
    if (ssplit) {
      Split<N>::at(g, iy, kz);
      jw = js = g;
    } else {
      iy = (int)(g / (N / 2 + 1));
      kz = (int)(g % (N / 2 + 1));
      js = g;
      jw = Split<N>::col(iy, kz);
    }
  }
};

// Rows t + m NT of column g in an x pass over an array whose kx planes lie
// `plane` elements apart (N (N/2+1) for the reference layout; the workspace
// and the driver's state copies use padded planes).
struct XRows {
  eidx base, step;
  __device__ __forceinline__ XRows(int t, int nt, long plane, long g)
      : base((eidx)t * plane + (eidx)g), step((eidx)nt * plane) {}
  __device__ __forceinline__ eidx operator()(int m) const { return base + (eidx)m * step; }
};

// A y-pass tile that lies in the Nyquist part of the workspace arrangement
// (whole tiles only: TK divides N N/2) on a grid of at most 128^3.  Its
// lines -- the Nyquist column of TK planes -- are contiguous, so its lanes
// run along them: on the small grids, whose passes last as long as their
// slowest tile, loads scattered over TK planes made that tile the slowest
// (64^3 FP64 y-inverse 0.038 against 0.024 ms).  On the larger grids those
// few tiles are hidden among thousands, and the column-major lanes keep their
// LDS exchanges free of bank conflicts.
template <int N, int TK>
__device__ __forceinline__ bool y_nyquist_tile(int xcd) {
  return N <= 128 && (Split<N>::MAIN % TK == 0) && tile_id(xcd) * TK >= Split<N>::MAIN;
}

// Rows of column q of a y-pass tile (q < N (N/2+1)) in the workspace
// arrangement: plane q / (N/2) at kz = q mod N/2 (q < N N/2, row stride N/2),
// or the Nyquist column of plane q - N N/2, whose ky values are contiguous
// (row stride 1).  Thread t of the column holds rows t + m NT.  When TK
// divides N N/2 a tile lies wholly in one of the two parts and the stride --
// taken from the tile's first column g0 -- is uniform over the block (scalar
// per-slot offsets); otherwise (small grids) it is per column.  body(at)
// runs the pass with at(m) = element offset of slot m in a field.  (One copy
// of the body: two copies, each with a compile-time stride, measured 1-3 %
// slower.)
template <int N, int TK, int NT, class Body>
__device__ __forceinline__ void y_rows(long q, long g0, long pw, int t, Body&& body) {
  constexpr int NR = N / 2;
  constexpr long MAIN = Split<N>::MAIN;
  const bool nyq = (MAIN % TK == 0) ? g0 >= MAIN : q >= MAIN;
  const long qq = (nyq && q < MAIN) ? MAIN : q;  // lanes past the last column stay in bounds
  const eidx b = nyq ? (eidx)(qq - MAIN) * pw + MAIN + t : (eidx)(q / NR) * pw + q % NR + (eidx)t * NR;
  const eidx st = nyq ? (eidx)NT : (eidx)NT * NR;
  body([=](int m) { return b + (eidx)m * st; });
}

// kz of column q of a y pass (N/2 in the Nyquist part)
template <int N>
__device__ __forceinline__ int y_kz(long q) { return q < Split<N>::MAIN ? (int)(q % (N / 2)) : N / 2; }

// ---------------------------------------------------------------------------
// 1. x-inverse
// ---------------------------------------------------------------------------
template <typename T, int N, int LES, class PL>
__global__ void __launch_bounds__((WideShape<T, N, PL>::THREADS))
    k_xinv(const cpx<T>* __restrict__ uh, long ps, int ssplit, cpx<T>* __restrict__ W, long pw,
           const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  uh = member_ptr(uh, mb.state);
  W = member_ptr(W, mb.work);
  using S = WideShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, +1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const XCol<N> xc(col.g, ssplit);
  const XRows rs(t, NT, ps, xc.js), rw(t, NT, pw, xc.jw);  // U_hat rows, W rows
  const long CS = N * ps, C = N * pw;
  cpx<T> v[P];
  if constexpr (LES) {
#pragma unroll 1
    for (int j = 0; j < 3; ++j) {
      cpx<T> w[P];
#pragma unroll
      for (int m = 0; m < P; ++m) {
        v[m] = ldx(uh + j * CS, rs(m));
        w[m] = mul_i((T)kfreq(t + m * NT, N) * v[m]);  // 1j * K0 * U_j
      }
      FFT::run(v, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, rw(m), v[m]);
      FFT::run(w, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + (3 + j) * C, rw(m), w[m]);
    }
  } else {
    // U_hat and the x-derivatives of its y and z components, from one read
    // of the three components: W0-2 = F_x^-1 U_j, W3 = F_x^-1 (1j Kx U_1),
    // W4 = F_x^-1 (1j Kx U_2).  The y pass completes omega_hat = 1j (K x U_hat)
    // (math_formula.py:22-24) from these five fields.
    cpx<T> u0[P], u1[P], u2[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
      u0[m] = ldx(uh, rs(m));
      u1[m] = ldx(uh + CS, rs(m));
      u2[m] = ldx(uh + 2 * CS, rs(m));
    }
    auto kx = [&](int m) { return (T)kfreq(t + m * NT, N); };
    auto fft_store = [&](cpx<T>(&x)[P], int j) {
      FFT::run(x, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, rw(m), x[m]);
    };
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * u1[m]);
    fft_store(v, 3);
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * u2[m]);
    fft_store(v, 4);
    fft_store(u0, 0);
    fft_store(u1, 1);
    fft_store(u2, 2);
  }
}

// ---------------------------------------------------------------------------
// 2. y-inverse (in place on W)
// ---------------------------------------------------------------------------
template <typename T, int N, int LES, class PL>
__global__ void __launch_bounds__((WideShape<T, N, PL>::THREADS))
    k_yinv(cpx<T>* W, long pw, const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  W = member_ptr(W, mb.work);
  using S = WideShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, +1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd, NT, y_nyquist_tile<N, TK>(xcd));
  const int t = col.t;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const long iz = y_kz<N>(col.g);
  const long C = N * pw;
  y_rows<N, TK, NT>(col.g, col.g0, pw, t, [&](auto at) {
  cpx<T> v[P];
  if constexpr (LES) {
#pragma unroll 1
    for (int j = 0; j < 3; ++j) {
      cpx<T> w[P];
#pragma unroll
      for (int m = 0; m < P; ++m) {
        v[m] = ldx(W + j * C, at(m));
        w[m] = mul_i((T)kfreq(t + m * NT, N) * v[m]);  // 1j * K1 * (.)
      }
      FFT::run(v, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, at(m), v[m]);
      FFT::run(w, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + (6 + j) * C, at(m), w[m]);
    }
  }
  if constexpr (LES) {
#pragma unroll 1
    for (int j = 3; j < 6; ++j) {
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = ldx(W + j * C, at(m));
      FFT::run(v, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, at(m), v[m]);
    }
  } else {
    // from W0-2 = U_j, W3 = Dx U_1, W4 = Dx U_2 (x already inverse-transformed):
    // omega = 1j (K x U) = (1j (Ky w - Kz v), 1j Kz u - Dx w, Dx v - 1j Ky u),
    // then every field y-transformed: W0-2 = U, W3-5 = omega
    auto ky = [&](int m) { return (T)kfreq(t + m * NT, N); };
    const T kz = (T)iz;
    auto fft_store = [&](cpx<T>(&x)[P], int j) {
      FFT::run(x, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, at(m), x[m]);
    };
    cpx<T> u[P], w[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
      u[m] = ldx(W, at(m));
      v[m] = ldx(W + 3 * C, at(m));  // Dx v
    }
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = v[m] - mul_i(ky(m) * u[m]);  // omega_z
    fft_store(v, 5);
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = mul_i(kz * u[m]) - ldx(W + 4 * C, at(m));  // omega_y
    fft_store(v, 4);
    fft_store(u, 0);
#pragma unroll
    for (int m = 0; m < P; ++m) {
      u[m] = ldx(W + C, at(m));      // v
      w[m] = ldx(W + 2 * C, at(m));  // w
      v[m] = mul_i(ky(m) * w[m] - kz * u[m]);  // omega_x
    }
    fft_store(v, 3);
    fft_store(u, 1);
    fft_store(w, 2);
  }
  });
}

// ---------------------------------------------------------------------------
// 3. z kernel: one line per (x, y) row.
// ---------------------------------------------------------------------------
// Lines of up to 64 threads (one wavefront or less): single-wavefront blocks
// of LPB = 64 / NT lines with wavefront-local exchanges (4 % faster than
// 256-thread blocks at 256^3 FP64, 3 % FP32).  Longer lines (radix-4 at 512
// points: 128 threads) need block barriers: one line per block, so a barrier
// waits for that line's two wavefronts only, and ping-pong exchange buffers
// (ZPPB), one barrier per exchange (512^3 FP64 z pass 6.28 -> 5.24 ms).
template <class PL>
struct ZLines {
  static constexpr int NT = PL::NT;
  static constexpr int LPB = NT > 64 ? 1 : ((64 / NT) > 0 ? (64 / NT) : 1);
  static constexpr int THREADS = LPB * NT;
};
template <typename T, int N, class PL>
constexpr int ZPPB = WaveLocal<PL>::value ? 0 : ZLines<PL>::LPB * LdsLine<T, N>::STRIDE;

// Line pairs Z_k = A_k + i B_k of two half-spectra A, B (uniform field bases,
// line at row0, Nyquist bins staged in LDS): pair_issue loads bins k < N/2 once,
// pair_finish completes the Hermitian half through LDS, multiplies by i*kz
// where asked and drops the imaginary parts of the kz = 0 and N/2 bins (as
// c2r does).
// Split in two so that the next pair's loads are issued before the current
// pair's transform.
// NQ: the line's Nyquist bin is staged in LDS (see z_row0) instead of held
// in registers by thread 0.
template <typename T, int P, bool NQ>
struct PairRaw {
  static constexpr int S = NQ ? P / 2 : P / 2 + 1;
  cpx<T> a[S], b[S];  // bins k = t + m*NT < N/2 (and k = N/2, thread 0, without NQ)
};

// A z line (plane x, row y) of the workspace arrangement: bins k < N/2 at
// x pw + y N/2 + k (row0 + k, 32-bit element offsets), the Nyquist bin
// k = N/2 in the plane's Nyquist column at x pw + N N/2 + y.  Only thread 0
// of a line holds that bin.  In FP32 the z kernels stage the Nyquist bins of
// their lines through LDS (loaded together at the start, stored together at
// the end), so that no second address and no extra slot per line stay in
// registers: the radix-8 kernels kept their 2 waves per SIMD that way (the
// direct form needed more than 256 VGPRs); in FP64 thread 0 loads and stores
// the bin directly (measured 6 % faster there: 256^3 z pass 0.502 against
// 0.534 ms, 512^3 4.67 against 4.95 ms).
template <int N>
__device__ __forceinline__ unsigned z_row0(long r, long pw) {
  return (unsigned)((r / N) * pw + (r % N) * (N / 2));
}
template <int N>
__device__ __forceinline__ unsigned z_nyquist(long r, long pw) {
  return (unsigned)((r / N) * pw + Split<N>::MAIN + r % N);
}

// Nyquist bins of fields f < NF (field stride C) for the block's LPB lines
// (rows row_base ...), destined for nq[f * LPB + line] (lines past the last
// row read row 0).
// When the block has at least NF LPB lanes, lane i issues the load of value
// i (every lane loads, branch-free; lanes past NF LPB read value 0 again) and
// nyquist_stage stores it later, so the load's latency overlaps whatever is
// issued in between; otherwise nyquist_issue loads them all into nq itself.
template <typename T, int N, int LPB, int NF, int THREADS>
__device__ __forceinline__ cpx<T> nyquist_issue(cpx<T>* nq, const cpx<T>* W, long C, long pw, long row_base) {
  auto value = [&](int i) {
    const int f = i / LPB, l = i % LPB;
    long r = row_base + l;
    if (r >= (long)N * N) r = 0;
    return ld32(W + f * C, z_nyquist<N>(r, pw));
  };
  if constexpr (NF * LPB <= THREADS) {
    return value(threadIdx.x < NF * LPB ? (int)threadIdx.x : 0);
  } else {
    for (int i = threadIdx.x; i < NF * LPB; i += THREADS) nq[i] = value(i);
    return cpx<T>{T(0), T(0)};
  }
}
// nq must have room for max(NF LPB, THREADS) values; ends with the line
// synchronisation
template <typename T, int LPB, int NF, int THREADS, bool WAVE>
__device__ __forceinline__ void nyquist_stage(cpx<T>* nq, cpx<T> v) {
  if constexpr (NF * LPB <= THREADS) nq[threadIdx.x] = v;  // lanes past NF LPB write unused slots
  line_sync<WAVE>();
}
// nq[f * LPB + line] -> the Nyquist bins of fields f < nf (valid lines)
template <typename T, int N, int LPB, bool WAVE>
__device__ __forceinline__ void nyquist_store(const cpx<T>* nq, cpx<T>* W, long C, int nf, long pw, long row_base) {
  line_sync<WAVE>();
  for (int i = threadIdx.x; i < nf * LPB; i += blockDim.x) {
    const int f = i / LPB, l = i % LPB;
    const long r = row_base + l;
    if (r < (long)N * N) st32(W + f * C, z_nyquist<N>(r, pw), nq[i]);
  }
}

template <typename T, int N, int P, int NT, bool NQ>
__device__ __forceinline__ void pair_issue(PairRaw<T, P, NQ>& r, const cpx<T>* A, const cpx<T>* B, unsigned row0,
                                           unsigned nyq, int t) {
#pragma unroll
  for (int m = 0; m < P / 2; ++m) {
    r.a[m] = ld32(A, row0 + t + m * NT);
    r.b[m] = ld32(B, row0 + t + m * NT);
  }
  if constexpr (!NQ) {
    if (t == 0) {
      r.a[P / 2] = ld32(A, nyq);
      r.b[P / 2] = ld32(B, nyq);
    }
  }
}
// Each half-spectrum element is loaded once; the completion k > N/2 takes
// conj of bin N - k from the line's LDS row (pmap: the unpermuted layout);
// the Nyquist bins nqa, nqb come from the LDS staging area.  The row must
// be free on entry (the previous transform ended with its barrier); it is
// free again on return.
template <typename T, int N, int P, int NT, bool WAVE, bool NQ, class Map>
__device__ __forceinline__ void pair_finish(cpx<T> (&v)[P], const PairRaw<T, P, NQ>& r, const cpx<T>& nqa,
                                            const cpx<T>& nqb, bool dzA, bool dzB, int t, cpx<T>* lds,
                                            const Map& pmap) {
  auto complete = [&](const cpx<T>(&h)[PairRaw<T, P, NQ>::S], const cpx<T>& nq, bool dz, bool second) {
#pragma unroll
    for (int m = 0; m < P / 2; ++m) lds[pmap(t + m * NT)] = h[m];
    if constexpr (NQ)
      lds[pmap(N / 2)] = nq;  // by every lane of the line (the same value): no divergent branch
    else if (t == 0)
      lds[pmap(N / 2)] = h[P / 2];
    line_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int k = t + m * NT;
      const int kk = k <= N / 2 ? k : N - k;
      cpx<T> x = lds[pmap(kk)];
      if (dz) x = mul_i((T)kk * x);
      if (k > N / 2)
        x = conj(x);
      else if (k == 0 || k == N / 2)
        x.y = T(0);
      v[m] = second ? cpx<T>{v[m].x - x.y, v[m].y + x.x} : x;  // a + i b
    }
    line_sync<WAVE>();
  };
  complete(r.a, nqa, dzA, false);
  complete(r.b, nqb, dzB, true);
}

// After a forward FFT of a + i b held in v (natural layout), split into the
// half-spectra A_k, B_k of a and b for this thread's bins k = t + m*NT <= N/2,
// m = 0..P/2.  Uses the line's LDS row through an unpermuted map (the
// line's LdsLine::plain()); ends with a barrier.
template <typename T, int N, int P, int NT, bool WAVE = false, class Map>
__device__ __forceinline__ void split_pair(const cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t,
                                           cpx<T> (&A)[P / 2 + 1], cpx<T> (&B)[P / 2 + 1]) {
#pragma unroll
  for (int m = 0; m < P; ++m) lds[map(t + m * NT)] = v[m];
  line_sync<WAVE>();
#pragma unroll
  for (int m = 0; m <= P / 2; ++m) {
    const int k = t + m * NT;
    const cpx<T> zk = v[m];
    const cpx<T> zn = conj(lds[map((N - k) % N)]);
    const cpx<T> d = zk - zn;
    A[m] = T(0.5) * (zk + zn);
    B[m] = {T(0.5) * d.y, T(-0.5) * d.x};
  }
  line_sync<WAVE>();
}

template <typename T>
__device__ __forceinline__ T nanmax_t(T a, T b) {
  return (isnan(a) || a > b) ? a : b;
}

__device__ __forceinline__ void block_max_store(double v, double* out) {
  __shared__ double s[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = (isnan(v) || v > w) ? v : w;
  }
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) s[wid] = v;
  block_sync();
  if (threadIdx.x == 0) {
    double r = s[0];
    for (int i = 1; i < (int)(blockDim.x + 63) / 64; ++i) r = (isnan(r) || r > s[i]) ? r : s[i];
    out[blockIdx.x] = r;
  }
}

template <typename T, int N, class PL, int MODEL>
__global__ void __launch_bounds__(ZLines<PL>::THREADS)
    k_zles(cpx<T>* W, long pw, T scale, T coeff, double* umax_part, const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  W = member_ptr(W, mb.work);
  if (umax_part) umax_part = member_ptr(umax_part, mb.work);
  if (mb.coeff) coeff = (T)mb.coeff[blockIdx.y];
  constexpr int P = PL::P, NT = PL::NT, NH = N / 2 + 1, KH = P / 2 + 1;
  constexpr int LPB = ZLines<PL>::LPB;
  constexpr bool WV = WaveLocal<PL>::value;  // each line inside one wavefront: no block barriers
  using IFFT = LineFFTP<T, PL, +1, WV, ZPPB<T, N, PL>>;
  using FFT = LineFFTP<T, PL, -1, WV, ZPPB<T, N, PL>>;
  __shared__ cpx<T> lds[LPB * LdsLine<T, N>::STRIDE + ZPPB<T, N, PL>];
  constexpr int TH = ZLines<PL>::THREADS;
  constexpr bool NQ = sizeof(T) == 4;  // Nyquist bins staged in LDS (z_row0)
  __shared__ cpx<T> nq[!NQ ? 1 : 9 * LPB > TH ? 9 * LPB : TH];  // Nyquist bins of the block's lines: 9 inputs, then 8 outputs
  const int line = threadIdx.x / NT, t = threadIdx.x % NT;
  // lines in XCD-aware block order: consecutive rows y of a plane -- whose
  // Nyquist bins share cache lines -- run in the same XCD's L2
  const long row_base = tile_id(xcd) * LPB;
  const long row = row_base + line;
  const bool valid = row < (long)N * N;
  const unsigned row0 = z_row0<N>(valid ? row : 0, pw);  // (x, y) line: plane x = row / N, row y = row % N
  const long C = N * pw;
  const LdsLine<T, N> map{line * LdsLine<T, N>::STRIDE};
  const TwRegs<T, PL> twr(tw, t);  // read once, used by all 11 transforms
  cpx<T>* w[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) w[i] = W + i * C;
  const unsigned nyq_addr = z_nyquist<N>(valid ? row : 0, pw);  // thread 0's Nyquist bin, without NQ
  auto nyq = [&](int f) -> cpx<T>& { return nq[NQ ? f * LPB + line : 0]; };
  // Nyquist bin of output field f (thread 0 of each line holds it)
  auto put_nyq = [&](int f, const cpx<T>& val) {
    if constexpr (NQ) {
      if (t == 0) nyq(f) = val;
    } else {
      if (valid && t == 0) st32(w[f], nyq_addr, val);
    }
  };

  // ---- inverse: 6 complex transforms -> 12 physical fields (scaled 1/N^3)
  T u0[P], u1[P], u2[P], a11[P], a12[P], a13[P], a21[P], a22[P], a23[P], a31[P], a32[P], a33[P];
  cpx<T> v[P];
  // pair i+1 is loaded while pair i is transformed; the Nyquist bins' loads
  // follow the first pair's, so their latencies overlap
  PairRaw<T, P, NQ> pr;
  cpx<T> nqv{};
  if constexpr (NQ) nqv = nyquist_issue<T, N, LPB, 9, TH>(nq, W, C, pw, row_base);
  pair_issue<T, N, P, NT>(pr, w[0], w[1], row0, nyq_addr, t);
  if constexpr (NQ) nyquist_stage<T, LPB, 9, TH, WV>(nq, nqv);
#define SLES_INV(FA, FB, DA, DB, NA, NB, RA, RB)                                     \
  pair_finish<T, N, P, NT, WV>(v, pr, nyq(FA), nyq(FB), DA, DB, t, lds, map.plain()); \
  if (NA >= 0) pair_issue<T, N, P, NT>(pr, w[NA], w[NB], row0, nyq_addr, t);           \
  IFFT::run(v, lds, map, t, twr);                                                    \
  _Pragma("unroll") for (int m = 0; m < P; ++m) {                                    \
    RA[m] = scale * v[m].x;                                                          \
    RB[m] = scale * v[m].y;                                                          \
  }
  SLES_INV(0, 1, false, false, 2, 0, u0, u1)
  SLES_INV(2, 0, false, true, 1, 2, u2, a31)
  SLES_INV(1, 2, true, true, 3, 4, a32, a33)
  SLES_INV(3, 4, false, false, 5, 6, a11, a12)
  SLES_INV(5, 6, false, false, 7, 8, a13, a21)
  SLES_INV(7, 8, false, false, -1, -1, a22, a23)
#undef SLES_INV

  // ---- pointwise: omega, u x omega, nu_R, tau = nu_R S_ij, max(u)
  T q0[P], q1[P], q2[P], q3[P], q4[P], q5[P], q6[P], q7[P], q8[P];
  T umax = -INFINITY;
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const T w0 = a23[m] - a32[m], w1 = a31[m] - a13[m], w2 = a12[m] - a21[m];
    q0[m] = u1[m] * w2 - u2[m] * w1;
    q1[m] = u2[m] * w0 - u0[m] * w2;
    q2[m] = u0[m] * w1 - u1[m] * w0;
    umax = nanmax_t(umax, nanmax_t(u0[m], nanmax_t(u1[m], u2[m])));
    T nu;
    const T s12 = a12[m] + a21[m], s13 = a13[m] + a31[m], s23 = a23[m] + a32[m];
    if constexpr (MODEL == SLES_LES_SMAGORINSKY) {
      const T ss = a11[m] * a11[m] + T(0.5) * (s12 * s12) + a22[m] * a22[m] + T(0.5) * (s13 * s13) +
                   T(0.5) * (s23 * s23) + a33[m] * a33[m];
      nu = sqrt(T(2) * ss) * coeff;
    } else {
      const T b11 = a11[m] * a11[m] + a21[m] * a21[m] + a31[m] * a31[m];
      const T b22 = a12[m] * a12[m] + a22[m] * a22[m] + a32[m] * a32[m];
      const T b33 = a13[m] * a13[m] + a23[m] * a23[m] + a33[m] * a33[m];
      const T b12 = a11[m] * a12[m] + a21[m] * a22[m] + a31[m] * a32[m];
      const T b13 = a11[m] * a13[m] + a21[m] * a23[m] + a31[m] * a33[m];
      const T b23 = a12[m] * a13[m] + a22[m] * a23[m] + a32[m] * a33[m];
      const T sb = sqrt(b11 * b22 - b12 * b12 + b11 * b33 - b13 * b13 + b22 * b33 - b23 * b23);
      const T aa = b11 + b22 + b33;
      const T f = (aa > T(1e-4) && sb > aa / T(1e7)) ? sqrt(T(1) / aa) : T(0);
      nu = sb * (f * coeff);
    }
    q3[m] = a11[m] * nu;
    q4[m] = T(0.5) * s12 * nu;
    q5[m] = T(0.5) * s13 * nu;
    q6[m] = a22[m] * nu;
    q7[m] = T(0.5) * s23 * nu;
    q8[m] = a33[m] * nu;
  }
  if (umax_part) block_max_store(valid ? (double)umax : -INFINITY, umax_part);

  // ---- forward: 5 complex transforms -> 8 output half-spectra; slots
  // m < P/2 (bins below N/2) go to memory, thread 0's Nyquist bin (slot P/2)
  // to the LDS staging area (outputs overwrite the inputs' slots, whose
  // values are no longer needed), stored with the others at the end
  cpx<T> A[KH], B[KH], t13[P / 2], t23[P / 2];
  constexpr T KN2 = T(N);  // 2 kz at the Nyquist bin
#define SLES_FWD(QA, QB)                                     \
  _Pragma("unroll") for (int m = 0; m < P; ++m) v[m] = {QA[m], QB[m]}; \
  FFT::run(v, lds, map, t, twr);                             \
  split_pair<T, N, P, NT, WV>(v, lds, map.plain(), t, A, B);
  SLES_FWD(q4, q5)  // tau_xy, tau_xz
#pragma unroll
  for (int m = 0; m < P / 2; ++m) {
    t13[m] = B[m];
    if (valid) {
      st32(w[4], row0 + t + m * NT, A[m]);
      st32(w[5], row0 + t + m * NT, B[m]);
    }
  }
  put_nyq(4, A[P / 2]);
  put_nyq(5, B[P / 2]);
  const cpx<T> n13 = B[P / 2];  // without NQ, thread 0 keeps it for N_x
  SLES_FWD(q6, q7)  // tau_yy, tau_yz
#pragma unroll
  for (int m = 0; m < P / 2; ++m) {
    t23[m] = B[m];
    if (valid) {
      st32(w[6], row0 + t + m * NT, A[m]);
      st32(w[7], row0 + t + m * NT, B[m]);
    }
  }
  put_nyq(6, A[P / 2]);
  put_nyq(7, B[P / 2]);
  const cpx<T> n23 = B[P / 2];
  SLES_FWD(q2, q8)  // N_z, tau_zz
#pragma unroll
  for (int m = 0; m < P / 2; ++m) {
    const T kz2 = T(2) * (T)(t + m * NT);
    if (valid) st32(w[2], row0 + t + m * NT, A[m] + mul_i(kz2 * B[m]));
  }
  put_nyq(2, A[P / 2] + mul_i(KN2 * B[P / 2]));
  SLES_FWD(q0, q1)  // N_x, N_y
#pragma unroll
  for (int m = 0; m < P / 2; ++m) {
    const T kz2 = T(2) * (T)(t + m * NT);
    if (valid) {
      st32(w[0], row0 + t + m * NT, A[m] + mul_i(kz2 * t13[m]));
      st32(w[1], row0 + t + m * NT, B[m] + mul_i(kz2 * t23[m]));
    }
  }
  put_nyq(0, A[P / 2] + mul_i(KN2 * (NQ ? nyq(5) : n13)));
  put_nyq(1, B[P / 2] + mul_i(KN2 * (NQ ? nyq(7) : n23)));
  T zq[P];
#pragma unroll
  for (int m = 0; m < P; ++m) zq[m] = T(0);
  SLES_FWD(q3, zq)  // tau_xx
#pragma unroll
  for (int m = 0; m < P / 2; ++m)
    if (valid) st32(w[3], row0 + t + m * NT, A[m]);
  put_nyq(3, A[P / 2]);
#undef SLES_FWD
  if constexpr (NQ) nyquist_store<T, N, LPB, WV>(nq, W, C, 8, pw, row_base);
}

template <typename T, int N, class PL>
__global__ void __launch_bounds__(ZLines<PL>::THREADS)
    k_zns(cpx<T>* W, long pw, T scale, double* umax_part, const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  W = member_ptr(W, mb.work);
  if (umax_part) umax_part = member_ptr(umax_part, mb.work);
  constexpr int P = PL::P, NT = PL::NT, NH = N / 2 + 1, KH = P / 2 + 1;
  constexpr int LPB = ZLines<PL>::LPB;
  constexpr bool WV = WaveLocal<PL>::value;  // each line inside one wavefront: no block barriers
  using IFFT = LineFFTP<T, PL, +1, WV, ZPPB<T, N, PL>>;
  using FFT = LineFFTP<T, PL, -1, WV, ZPPB<T, N, PL>>;
  __shared__ cpx<T> lds[LPB * LdsLine<T, N>::STRIDE + ZPPB<T, N, PL>];
  constexpr int TH = ZLines<PL>::THREADS;
  constexpr bool NQ = sizeof(T) == 4;  // Nyquist bins staged in LDS (z_row0)
  __shared__ cpx<T> nq[!NQ ? 1 : 6 * LPB > TH ? 6 * LPB : TH];  // Nyquist bins of the block's lines: 6 inputs, then 3 outputs
  const int line = threadIdx.x / NT, t = threadIdx.x % NT;
  // lines in XCD-aware block order: consecutive rows y of a plane -- whose
  // Nyquist bins share cache lines -- run in the same XCD's L2
  const long row_base = tile_id(xcd) * LPB;
  const long row = row_base + line;
  const bool valid = row < (long)N * N;
  const unsigned row0 = z_row0<N>(valid ? row : 0, pw);  // (x, y) line: plane x = row / N, row y = row % N
  const long C = N * pw;
  const LdsLine<T, N> map{line * LdsLine<T, N>::STRIDE};
  const TwRegs<T, PL> twr(tw, t);  // read once, used by all 5 transforms
  cpx<T>* w[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) w[i] = W + i * C;
  const unsigned nyq_addr = z_nyquist<N>(valid ? row : 0, pw);  // thread 0's Nyquist bin, without NQ
  auto nyq = [&](int f) -> cpx<T>& { return nq[NQ ? f * LPB + line : 0]; };
  // Nyquist bin of output field f (thread 0 of each line holds it)
  auto put_nyq = [&](int f, const cpx<T>& val) {
    if constexpr (NQ) {
      if (t == 0) nyq(f) = val;
    } else {
      if (valid && t == 0) st32(w[f], nyq_addr, val);
    }
  };
  T u0[P], u1[P], u2[P], o0[P], o1[P], o2[P];
  cpx<T> v[P];
  // pair i+1 is loaded while pair i is transformed; the Nyquist bins' loads
  // follow the first pair's, so their latencies overlap
  PairRaw<T, P, NQ> pr;
  cpx<T> nqv{};
  if constexpr (NQ) nqv = nyquist_issue<T, N, LPB, 6, TH>(nq, W, C, pw, row_base);
  pair_issue<T, N, P, NT>(pr, w[0], w[1], row0, nyq_addr, t);
  if constexpr (NQ) nyquist_stage<T, LPB, 6, TH, WV>(nq, nqv);
#define SLES_INV(FA, NA, RA, RB)                                                            \
  pair_finish<T, N, P, NT, WV>(v, pr, nyq(FA), nyq(FA + 1), false, false, t, lds, map.plain()); \
  if (NA >= 0) pair_issue<T, N, P, NT>(pr, w[NA], w[NA + 1], row0, nyq_addr, t);              \
  IFFT::run(v, lds, map, t, twr);                                                           \
  _Pragma("unroll") for (int m = 0; m < P; ++m) {                                           \
    RA[m] = scale * v[m].x;                                                                 \
    RB[m] = scale * v[m].y;                                                                 \
  }
  SLES_INV(0, 2, u0, u1)
  SLES_INV(2, 4, u2, o0)
  SLES_INV(4, -1, o1, o2)
#undef SLES_INV
  T umax = -INFINITY;
  T n0[P], n1[P], n2[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    n0[m] = u1[m] * o2[m] - u2[m] * o1[m];
    n1[m] = u2[m] * o0[m] - u0[m] * o2[m];
    n2[m] = u0[m] * o1[m] - u1[m] * o0[m];
    umax = nanmax_t(umax, nanmax_t(u0[m], nanmax_t(u1[m], u2[m])));
  }
  if (umax_part) block_max_store(valid ? (double)umax : -INFINITY, umax_part);
  cpx<T> A[KH], B[KH];
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = {n0[m], n1[m]};
  FFT::run(v, lds, map, t, twr);
  split_pair<T, N, P, NT, WV>(v, lds, map.plain(), t, A, B);
#pragma unroll
  for (int m = 0; m < P / 2; ++m)
    if (valid) {
      st32(w[0], row0 + t + m * NT, A[m]);
      st32(w[1], row0 + t + m * NT, B[m]);
    }
  put_nyq(0, A[P / 2]);
  put_nyq(1, B[P / 2]);
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = {n2[m], T(0)};
  FFT::run(v, lds, map, t, twr);
  split_pair<T, N, P, NT, WV>(v, lds, map.plain(), t, A, B);
#pragma unroll
  for (int m = 0; m < P / 2; ++m)
    if (valid) st32(w[2], row0 + t + m * NT, A[m]);
  put_nyq(2, A[P / 2]);
  if constexpr (NQ) nyquist_store<T, N, LPB, WV>(nq, W, C, 3, pw, row_base);
}

// ---------------------------------------------------------------------------
// 4. y-forward (in place)
// ---------------------------------------------------------------------------
template <typename T, int N, int LES, class PL>
__global__ void __launch_bounds__((WideShape<T, N, PL>::THREADS))
    k_yfwd(cpx<T>* W, long pw, const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  W = member_ptr(W, mb.work);
  using S = WideShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  // ping-pong exchange buffers in FP32 only: in FP64 they would double the
  // block's LDS to 128 KiB, which measured slower
  constexpr int PPB = sizeof(T) == 4 ? N * TK : 0;
  using FFT = LineFFTP<T, PL, -1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd, NT, y_nyquist_tile<N, TK>(xcd));
  const int t = col.t;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const long C = N * pw;
  y_rows<N, TK, NT>(col.g, col.g0, pw, t, [&](auto at) {
  cpx<T> v[P], acc[P];
  auto load = [&](int f) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = ldx(W + f * C, at(m));
  };
  auto store = [&](int f, const cpx<T>(&x)[P]) {
    if (col.valid)
#pragma unroll
      for (int m = 0; m < P; ++m) stx(W + f * C, at(m), x[m]);
  };
  if constexpr (LES) {
    // H_i = F_y(G_i) + 2i Ky F_y(tau_iy): tau_xy (W4), tau_yy (W6), tau_yz (W7)
    // fields in the order 4, 0, 6, 1, 7, 2, 3, 5; the next field's loads are
    // issued before the current field's transform
    auto field = [](int q) {
      return q == 0 ? 4 : q == 1 ? 0 : q == 2 ? 6 : q == 3 ? 1 : q == 4 ? 7 : q == 5 ? 2 : q == 6 ? 3 : 5;
    };
    cpx<T> nx[P];
#pragma unroll
    for (int m = 0; m < P; ++m) nx[m] = ldx(W + 4 * C, at(m));
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = nx[m];
      if (q + 1 < 8) {
        const int fn = field(q + 1);
#pragma unroll
        for (int m = 0; m < P; ++m) nx[m] = ldx(W + fn * C, at(m));
      }
      FFT::run(v, lds, map, t, twr);
      if (q >= 6) {
        store(field(q), v);
      } else if ((q & 1) == 0) {
        if (q == 0) store(4, v);
#pragma unroll
        for (int m = 0; m < P; ++m) acc[m] = mul_i((T(2) * (T)kfreq(t + m * NT, N)) * v[m]);
      } else {
#pragma unroll
        for (int m = 0; m < P; ++m) acc[m] = v[m] + acc[m];
        store(q >> 1, acc);
      }
    }
  } else {
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
      load(f);
      FFT::run(v, lds, map, t, twr);
      store(f, v);
    }
  }
  });
}

// ---------------------------------------------------------------------------
// 5. x-forward + projection + scheme epilogue
// ---------------------------------------------------------------------------
__device__ __forceinline__ double phi1_d(double z) {
  if (fabs(z) < 1e-2) return 1 + z / 2 * (1 + z / 3 * (1 + z / 4 * (1 + z / 5 * (1 + z / 6))));
  return expm1(z) / z;
}
__device__ __forceinline__ double phi2_d(double z) {
  if (fabs(z) < 1e-2) return 0.5 * (1 + z / 3 * (1 + z / 4 * (1 + z / 5 * (1 + z / 6 * (1 + z / 7)))));
  return (expm1(z) - z) / (z * z);
}

// ynext (RK4 stages 0-2): also receives the next stage's input U_hat.  e and
// C: element and field stride of the RK-state arrays (and of a.stages).
template <typename T>
__device__ __forceinline__ void epilogue(cpx<T> (&s)[3], eidx e, int kx, int ky, int kz, int n, const StepArgs& a,
                                         double dt, cpx<T>* uh, cpx<T>* uh0, cpx<T>* uh1, cpx<T>* aux, long C,
                                         cpx<T> (&ynext)[3]) {
  const KIndex k = {kx, ky, kz};
  const T keep = dealias_keep(k, n) ? T(1) : T(0);
  const long k2 = (long)kx * kx + (long)ky * ky + (long)kz * kz;
  const T den = (T)(k2 == 0 ? 1 : k2);
  const T kk[3] = {(T)kx, (T)ky, (T)kz};
  cpx<T> d[3];
  cpx<T> p = {T(0), T(0)};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    d[i] = keep * s[i];
    p = p + (kk[i] / den) * d[i];
  }
  cpx<T> u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    u[i] = ldx(uh + i * C, e);
    d[i] = d[i] - kk[i] * p;
    if (a.forcing > 0.0 && k2 != 0) d[i] = d[i] + (T)a.forcing * u[i];
  }
  if (a.scheme == EPI_NONLIN) {
#pragma unroll
    for (int i = 0; i < 3; ++i) stx(aux + i * C, e, d[i]);
    return;
  }
  if (a.scheme == EPI_ETD2) {
    // Cox-Matthews ETD2, L = -nu K^2, phi functions in FP64
    const double z = -dt * a.nu * (double)k2;
    const double p1 = phi1_d(z), p2 = phi2_d(z);
    const T ez = (T)exp(z), w1 = (T)(dt * (p1 + p2)), w0 = (T)(dt * p2);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const cpx<T> f0 = ldx(aux + i * C, e);
      stx(uh + i * C, e, (ez * u[i] + w1 * d[i]) - w0 * f0);
      stx(aux + i * C, e, d[i]);
    }
    return;
  }
  if (a.scheme == EPI_RK4 && a.stage == 0 && aux) {
#pragma unroll
    for (int i = 0; i < 3; ++i) stx(aux + i * C, e, d[i]);  // F(u_0), kept for an ETD2 start
  }
  if (a.nu > 0.0) {
    const T c = (T)(a.nu * (double)k2);
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] = d[i] - c * u[i];
  }
  if (a.scheme == EPI_RHS) {
#pragma unroll
    for (int i = 0; i < 3; ++i) stx(aux + i * C, e, d[i]);
    return;
  }
  if (a.scheme == EPI_RK4) {
    const T ca = (T)(a.ca * dt), cb = (T)(a.cb * dt);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cpx<T>* const ui = uh + i * C;
      cpx<T>* const u0i = uh0 + i * C;
      cpx<T>* const u1i = uh1 + i * C;
      if (a.stage == 0) {
        const cpx<T> y = u[i] + cb * d[i];
        stx(u0i, e, u[i]);
        stx(u1i, e, u[i] + ca * d[i]);
        stx(ui, e, y);
        if (a.stages) stx((cpx<T>*)a.stages + i * C, e, y);
        ynext[i] = y;
      } else if (a.stage < 3) {
        const cpx<T> y = ldx(u0i, e) + cb * d[i];
        stx(u1i, e, ldx(u1i, e) + ca * d[i]);
        stx(ui, e, y);
        if (a.stages) stx((cpx<T>*)a.stages + (3 * a.stage + i) * C, e, y);
        ynext[i] = y;
      } else {
        const cpx<T> fin = ldx(u1i, e) + ca * d[i];
        stx(u1i, e, fin);
        stx(ui, e, fin);
        ynext[i] = fin;  // for the physical-output x-inverse
      }
    }
    return;
  }
  // RK3 as written (src/RK3_method.py:27-33)
  const T hdt = (T)(0.5 * dt), fdt = (T)dt, tdt = (T)(2.0 * dt);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    cpx<T>* const ui = uh + i * C;
    cpx<T>* const u0i = uh0 + i * C;
    if (a.stage == 0) {
      stx(u0i, e, u[i]);
      stx(ui, e, u[i] + hdt * d[i]);
    } else if (a.stage == 1) {
      stx(ui, e, (ldx(u0i, e) - fdt * d[i]) + tdt * d[i]);
    } else {
      const cpx<T> sm = ((T(1) / T(6)) * d[i] + (T(2) / T(3)) * d[i]) + (T(1) / T(6)) * d[i];
      stx(ui, e, ldx(u0i, e) + fdt * sm);
    }
  }
}

// Blocks of at most 512 threads, except FP32 radix-4 (1024): the three
// accumulated components, the transform and the epilogue need ~200 VGPRs per
// thread with the radix-4 FP64 plan, which a 1024-thread block (128 VGPRs)
// would spill; the FP32 radix-4 form needs ~110 and gains 128-byte row
// segments per tile from the wider block (256^3 -5 %, 512^3 -10 %).
template <typename T, int N, class PL>
using XShape = StridedShape<T, N, PL, (sizeof(T) == 4 && PL::P == 4) ? 1024 : 512>;

// FUSE (RK4 stages 0-2, FP32 form): after the update, the block also runs
// the next stage's x-inverse on the new stage input it holds (same x-lines,
// same columns: no other block touches them), writing W as k_xinv would, so
// the next stage skips k_xinv and the re-read of U_hat.  FP32 256^3: 0.479 ms
// per fused launch against 0.372 + 0.146 ms separately (step +2 %).  In FP64
// the fused inverse runs on the x-forward's radix-4 plan at 2 waves per SIMD:
// over planes a multiple of 4 KiB apart it was slower than the separate
// radix-8 k_xinv (256^3: 1.071 against 1.038 ms); over padded planes it saves
// 0.1 ms of the 0.8 ms that the two kernels take.
// FLES: layout of the fused x-inverse output (LES stage layout, or 0: the
// layout without LES, which is also the physical-output pass's: the last RK4
// stage runs that pass's x-inverse on the new U_hat).
template <typename T, int N, int LES, class PL, bool FUSE = false, int FLES = LES>
__global__ void __launch_bounds__((XShape<T, N, PL>::THREADS))
    k_xfwd(cpx<T>* __restrict__ W, long pw, cpx<T>* uh, cpx<T>* uh0, cpx<T>* uh1, cpx<T>* aux, long ps, int ssplit,
           StepArgs a, const cpx<T>* __restrict__ tw, int xcd, Members mb) {
  W = member_ptr(W, mb.work);
  uh = member_ptr(uh, mb.state);
  if (uh0) uh0 = member_ptr(uh0, mb.state);
  if (uh1) uh1 = member_ptr(uh1, mb.state);
  if (aux) aux = member_ptr(aux, mb.state);
  if (a.dt_dev) a.dt_dev = member_ptr(a.dt_dev, mb.drive);
  using S = XShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, -1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  double dt = a.dt;
  bool skip = false;  // the step is not taken (device dt NaN or too small)
  if (a.dt_dev) {
    dt = *a.dt_dev;
    if (!(dt >= 1e-6)) {
      if constexpr (!FUSE) return;
      skip = true;  // fused: U_hat stays; the x-inverse below still runs on it
    }
  }
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const XCol<N> xc(col.g, ssplit);
  const XRows rw(t, NT, pw, xc.jw), rs(t, NT, ps, xc.js);  // W rows, RK-state rows
  const long C = N * pw, CS = N * ps;
  // a0..a2: S_i = F_x(W_i) (+ 2i Kx F_x(tau_xi) with LES).  Named arrays and
  // a rolled field loop with uniform branches keep them in registers.
  cpx<T> a0[P], a1[P], a2[P], v[P];
  // next field's loads are issued before the current field's transform
  constexpr int NF = LES ? 6 : 3;
  const int nf = skip ? 0 : NF;
  cpx<T> nx[P];
#pragma unroll
  for (int m = 0; m < P; ++m) nx[m] = ldx(W, rw(m));
#pragma unroll 1
  for (int f = 0; f < nf; ++f) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = nx[m];
    if (f + 1 < NF) {
#pragma unroll
      for (int m = 0; m < P; ++m) nx[m] = ldx(W + (f + 1) * C, rw(m));
    }
    FFT::run(v, lds, map, t, twr);
    if (f < 3) {
      if (f == 0) {
#pragma unroll
        for (int m = 0; m < P; ++m) a0[m] = v[m];
      } else if (f == 1) {
#pragma unroll
        for (int m = 0; m < P; ++m) a1[m] = v[m];
      } else {
#pragma unroll
        for (int m = 0; m < P; ++m) a2[m] = v[m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i((T(2) * (T)kfreq(t + m * NT, N)) * v[m]);
      if (f == 3) {
#pragma unroll
        for (int m = 0; m < P; ++m) a0[m] = a0[m] + v[m];
      } else if (f == 4) {
#pragma unroll
        for (int m = 0; m < P; ++m) a1[m] = a1[m] + v[m];
      } else {
#pragma unroll
        for (int m = 0; m < P; ++m) a2[m] = a2[m] + v[m];
      }
    }
  }
  const int ky = kfreq(xc.iy, N), kz = xc.kz;
  cpx<T> y0[P], y1[P], y2[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    cpx<T> s3[3] = {a0[m], a1[m], a2[m]};
    cpx<T> yn[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
    if (col.valid && !skip) {
      epilogue<T>(s3, rs(m), kfreq(t + m * NT, N), ky, kz, N, a, dt, uh, uh0, uh1, aux, CS, yn);
    } else if (col.valid) {
#pragma unroll
      for (int i = 0; i < 3; ++i) yn[i] = ldx(uh + i * CS, rs(m));
    }
    y0[m] = yn[0];
    y1[m] = yn[1];
    y2[m] = yn[2];
  }
  if constexpr (FUSE) {
    using IFFT = LineFFTP<T, PL, +1, false, PPB>;
    auto kx = [&](int m) { return (T)kfreq(t + m * NT, N); };
    auto ifft_store = [&](cpx<T>(&x)[P], int j) {
      IFFT::run(x, lds, map, t, twr);
      if (col.valid)
#pragma unroll
        for (int m = 0; m < P; ++m) stx(W + j * C, rw(m), x[m]);
    };
    if constexpr (FLES) {
      // as k_xinv: W_j = F_x^-1 Y_j, W_{3+j} = F_x^-1 (1j Kx Y_j)
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * y0[m]);
      ifft_store(v, 3);
      ifft_store(y0, 0);
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * y1[m]);
      ifft_store(v, 4);
      ifft_store(y1, 1);
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * y2[m]);
      ifft_store(v, 5);
      ifft_store(y2, 2);
    } else {
      // as k_xinv without LES: W0-2 = F_x^-1 Y, W3 = F_x^-1 (1j Kx Y_1), W4 = F_x^-1 (1j Kx Y_2)
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * y1[m]);
      ifft_store(v, 3);
#pragma unroll
      for (int m = 0; m < P; ++m) v[m] = mul_i(kx(m) * y2[m]);
      ifft_store(v, 4);
      ifft_store(y0, 0);
      ifft_store(y1, 1);
      ifft_store(y2, 2);
    }
  }
}

// First level of the CFL reduction: block b folds part[b*chunk, (b+1)*chunk)
// (NaN-propagating max) into out[b].
static __global__ void k_nanmax_chunks(const double* part, int np, int chunk, double* out, Members mb) {
  part = member_ptr(part, mb.work);
  out = member_ptr(out, mb.work);
  double m = -INFINITY;
  const int lo = blockIdx.x * chunk, hi = lo + chunk < np ? lo + chunk : np;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) m = (isnan(m) || m > part[i]) ? m : part[i];
  __shared__ double s[256];
  s[threadIdx.x] = m;
  block_sync();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double a = s[threadIdx.x], b = s[threadIdx.x + o];
      s[threadIdx.x] = (isnan(a) || a > b) ? a : b;
    }
    block_sync();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// dt = min(CFL dx / max(u), 0.1 dx^2 / nu)   (src/SM_solver.py:245-272)
static __global__ void k_cfl(const double* part, int np, double cfl, double dx, double nu, double* scalars) {
  double m = -INFINITY;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = (isnan(m) || m > part[i]) ? m : part[i];
  __shared__ double s[1024];
  s[threadIdx.x] = m;
  block_sync();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double a = s[threadIdx.x], b = s[threadIdx.x + o];
      s[threadIdx.x] = (isnan(a) || a > b) ? a : b;
    }
    block_sync();
  }
  if (threadIdx.x == 0) {
    const double umax = s[0];
    const double dt1 = cfl * dx / umax;
    const double dt2 = nu > 0 ? 0.1 * dx * dx / nu : 1e6;
    scalars[0] = umax;
    scalars[1] = (dt2 < dt1) ? dt2 : dt1;
  }
}

// The driver loop's per-step rules on the device (sles_step_drive; state
// layout DRV_* in sles.h): dt from the CFL rule on max(u) of the state about
// to be stepped (cfl_first until the first step is taken,
// src/SM_solver.py:21-23,60-63), then the checks the reference loop makes
// between two steps -- dt NaN ends the run (:64-67), dt < 1e-6 is the exit
// of compute_dt (:273-276); t starts at the first dt (:25); a loop top with
// t >= t_pause stops for the host's outputs (:40-45) unless the host resumed
// this loop top; t >= t_final or dt < 1e-8 ends the run (:46-47) -- and
// t += dt (:51).  state[DRV_DT_APPLY] receives the dt of the step, or NaN
// when it is not taken (the x-forward kernels then leave U_hat unchanged).
static __global__ void k_cfl_drive(const double* part, int np, double cfl, double cfl_first, double dx, double nu,
                                   double* st, Members mb) {
  part = member_ptr(part, mb.work);
  st = member_ptr(st, mb.drive);
  double m = -INFINITY;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = (isnan(m) || m > part[i]) ? m : part[i];
  __shared__ double s[1024];
  s[threadIdx.x] = m;
  block_sync();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double a = s[threadIdx.x], b = s[threadIdx.x + o];
      s[threadIdx.x] = (isnan(a) || a > b) ? a : b;
    }
    block_sync();
  }
  if (threadIdx.x != 0) return;
  const double umax = s[0];
  st[DRV_UMAX] = umax;
  const double status = st[DRV_STATUS];
  const bool resume = status == (double)SLES_DRIVE_RESUME;
  if (status != (double)SLES_DRIVE_RUNNING && !resume) {  // stopped earlier in this chunk
    st[DRV_DT_APPLY] = NAN;
    return;
  }
  const double c = st[DRV_STEPS] == 0.0 ? cfl_first : cfl;
  const double dt1 = c * dx / umax;
  const double dt2 = nu > 0 ? 0.1 * dx * dx / nu : 1e6;
  const double dt = (dt2 < dt1) ? dt2 : dt1;
  st[DRV_DT] = dt;
  double stop = 0.0;
  double t = st[DRV_T];
  if (isnan(dt)) {
    stop = SLES_DRIVE_NAN;
  } else if (dt < 1e-6) {
    stop = SLES_DRIVE_DT_TOO_SMALL;
  } else {
    if (isnan(t)) t = dt;
    st[DRV_T] = t;
    if (!resume && t >= st[DRV_T_PAUSE])
      stop = SLES_DRIVE_PAUSED;
    else if (t >= st[DRV_T_FINAL] || dt < 1e-8)
      stop = SLES_DRIVE_DONE;
  }
  if (stop != 0.0) {
    st[DRV_STATUS] = stop;
    st[DRV_DT_APPLY] = NAN;
    return;
  }
  st[DRV_T] = t + dt;
  st[DRV_STEPS] += 1.0;
  st[DRV_STATUS] = SLES_DRIVE_RUNNING;
  st[DRV_DT_APPLY] = dt;
}

// ---------------------------------------------------------------------------
// Optional per-kernel timing (measurement only).
// ---------------------------------------------------------------------------
enum { PK_XINV = 0, PK_YINV = 1, PK_Z = 2, PK_YFWD = 3, PK_XFWD = 4, PK_CFL = 5, PK_PHYS = 6, PK_N = 7 };
struct ProfRec {
  int kind;
  hipEvent_t a, b;
};
extern bool g_prof;
extern thread_local bool g_prof_off;  // set while sles_step_noalloc runs: no timing events
extern std::vector<ProfRec> g_recs;
extern std::vector<hipEvent_t> g_pool;

inline hipEvent_t prof_event() {
  if (g_pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  hipEvent_t e = g_pool.back();
  g_pool.pop_back();
  return e;
}

struct ProfScope {
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(int k, hipStream_t st) : kind(k), s(st) {
    if (g_prof && !g_prof_off && (a = prof_event())) (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = prof_event();
    if (!b) return;
    (void)hipEventRecord(b, s);
    g_recs.push_back({kind, a, b});
  }
};

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
template <typename T, int N, class PS, class PX, class PZ>
struct Fused {
  using SS = WideShape<T, N, PS>;  // x/y inverse and y forward
  using SX = XShape<T, N, PX>;  // x forward + stage update
  static constexpr int ZB = (int)(((long)N * N + ZLines<PZ>::LPB - 1) / ZLines<PZ>::LPB);
  // RK4 x-forward can run the next stage's x-inverse (see k_xfwd), in both
  // forms and at every size; the plan knob tune_fuse_x chooses per grid
  // (64^3 Taylor-Green 0.323 -> 0.312 ms per step; 256^3 FP64 Smagorinsky
  // through the driver loop over the padded workspace 8.49 -> 8.30 ms)
  static constexpr bool FUSE_X = true;

  // skip_xinv: W0-5 already hold the x-inverse output (fused into the previous x-forward)
  // nm members (grid rows) at the strides of mb; ps: kx-plane stride of the RK-state arrays
  static int stage(const sles_plan_s* p, int les, T coeff, cpx<T>* uh, long ps, int ssplit, cpx<T>* W,
                   double* part, hipStream_t s, bool skip_xinv, const Members& mb, int nm) {
    const long pw = plane_w(p);
    const cpx<T>* tw = (const cpx<T>*)p->tw;
    const int xcd = p->tune_xcd;
    constexpr long NCOL = (long)N * (N / 2 + 1);  // columns of every strided pass
    const dim3 gs((unsigned)((NCOL + SS::TK - 1) / SS::TK), (unsigned)nm), bs(SS::THREADS);
    const dim3 gz((unsigned)ZB, (unsigned)nm), bz(ZLines<PZ>::THREADS);
    const T scale = T(1.0 / ((double)N * N * N));
    if (les) {
      if (!skip_xinv) {
        ProfScope ps_(PK_XINV, s);
        hipLaunchKernelGGL((k_xinv<T, N, 1, PS>), gs, bs, 0, s, uh, ps, ssplit, W, pw, tw, xcd, mb);
      }
      { ProfScope ps_(PK_YINV, s); hipLaunchKernelGGL((k_yinv<T, N, 1, PS>), gs, bs, 0, s, W, pw, tw, xcd, mb); }
      ProfScope ps_(PK_Z, s);
      if (les == SLES_LES_SMAGORINSKY)
        hipLaunchKernelGGL((k_zles<T, N, PZ, SLES_LES_SMAGORINSKY>), gz, bz, 0, s, W, pw, scale, coeff, part, tw, xcd,
                           mb);
      else
        hipLaunchKernelGGL((k_zles<T, N, PZ, SLES_LES_VREMAN>), gz, bz, 0, s, W, pw, scale, coeff, part, tw, xcd,
                           mb);
    } else {
      if (!skip_xinv) {
        ProfScope ps_(PK_XINV, s);
        hipLaunchKernelGGL((k_xinv<T, N, 0, PS>), gs, bs, 0, s, uh, ps, ssplit, W, pw, tw, xcd, mb);
      }
      { ProfScope ps_(PK_YINV, s); hipLaunchKernelGGL((k_yinv<T, N, 0, PS>), gs, bs, 0, s, W, pw, tw, xcd, mb); }
      ProfScope ps_(PK_Z, s);
      hipLaunchKernelGGL((k_zns<T, N, PZ>), gz, bz, 0, s, W, pw, scale, part, tw, xcd, mb);
    }
    return check_launch("fused stage (inverse)") ? -1 : 0;
  }

  // fuse_next: also run the next stage's x-inverse (RK4 stages 0-2)
  // fuse_phys (last RK4 stage): also run the physical-output pass's x-inverse
  static int finish(const sles_plan_s* p, int les, cpx<T>* uh, cpx<T>* uh0, cpx<T>* uh1, cpx<T>* aux, long ps,
                    int ssplit, cpx<T>* W, const StepArgs& a, hipStream_t s, bool fuse_next, bool fuse_phys,
                    const Members& mb, int nm) {
    const long pw = plane_w(p);
    const cpx<T>* tw = (const cpx<T>*)p->tw;
    const int xcd = p->tune_xcd;
    constexpr long NCOL = (long)N * (N / 2 + 1);  // columns of every strided pass
    const dim3 gs((unsigned)((NCOL + SS::TK - 1) / SS::TK), (unsigned)nm), bs(SS::THREADS);
    const dim3 gx((unsigned)((NCOL + SX::TK - 1) / SX::TK), (unsigned)nm), bx(SX::THREADS);
    if (les) {
      { ProfScope ps_(PK_YFWD, s); hipLaunchKernelGGL((k_yfwd<T, N, 1, PS>), gs, bs, 0, s, W, pw, tw, xcd, mb); }
      ProfScope ps_(PK_XFWD, s);
      if constexpr (FUSE_X) {
        if (fuse_next) {
          hipLaunchKernelGGL((k_xfwd<T, N, 1, PX, true>), gx, bx, 0, s, W, pw, uh, uh0, uh1, aux, ps, ssplit, a, tw, xcd, mb);
          return check_launch("fused stage (forward)");
        }
        if (fuse_phys) {
          hipLaunchKernelGGL((k_xfwd<T, N, 1, PX, true, 0>), gx, bx, 0, s, W, pw, uh, uh0, uh1, aux, ps, ssplit, a, tw, xcd,
                             mb);
          return check_launch("fused stage (forward)");
        }
      }
      hipLaunchKernelGGL((k_xfwd<T, N, 1, PX>), gx, bx, 0, s, W, pw, uh, uh0, uh1, aux, ps, ssplit, a, tw, xcd, mb);
    } else {
      { ProfScope ps_(PK_YFWD, s); hipLaunchKernelGGL((k_yfwd<T, N, 0, PS>), gs, bs, 0, s, W, pw, tw, xcd, mb); }
      ProfScope ps_(PK_XFWD, s);
      if constexpr (FUSE_X)
        if (fuse_next || fuse_phys) {  // the same output layout
          hipLaunchKernelGGL((k_xfwd<T, N, 0, PX, true>), gx, bx, 0, s, W, pw, uh, uh0, uh1, aux, ps, ssplit, a, tw, xcd, mb);
          return check_launch("fused stage (forward)");
        }
      hipLaunchKernelGGL((k_xfwd<T, N, 0, PX>), gx, bx, 0, s, W, pw, uh, uh0, uh1, aux, ps, ssplit, a, tw, xcd, mb);
    }
    return check_launch("fused stage (forward)");
  }

  // skip_xinv: W0-4 already hold the x-inverse output (fused into the last x-forward)
  static int physical(const sles_plan_s* p, cpx<T>* uh, long ps, int ssplit, cpx<T>* W, void* u_out,
                      void* curl_out, hipStream_t s, bool skip_xinv = false) {
    const long pw = plane_w(p);
    const cpx<T>* tw = (const cpx<T>*)p->tw;
    const int xcd = p->tune_xcd;
    constexpr long NCOL = (long)N * (N / 2 + 1);  // columns of every strided pass
    const dim3 gs((unsigned)((NCOL + SS::TK - 1) / SS::TK)), bs(SS::THREADS);
    ProfScope ps_(PK_PHYS, s);
    if (!skip_xinv) hipLaunchKernelGGL((k_xinv<T, N, 0, PS>), gs, bs, 0, s, uh, ps, ssplit, W, pw, tw, xcd, Members{});
    hipLaunchKernelGGL((k_yinv<T, N, 0, PS>), gs, bs, 0, s, W, pw, tw, xcd, Members{});
    if (check_launch("fused physical")) return -1;
    const double sc = 1.0 / ((double)N * N * N);
    if (u_out && z_c2r_rows(p, W, pw, 1, u_out, 3L * N * N, sc, s)) return -1;
    if (curl_out && z_c2r_rows(p, W + 3 * N * pw, pw, 1, curl_out, 3L * N * N, sc, s)) return -1;
    return 0;
  }
};

// One fused step (or RHS) of `scheme` over the RK-state arrays uh, uh0, uh1,
// aux, whose kx planes lie ps elements apart and are arranged as the
// reference's (ssplit = 0, ps = N (N/2+1)) or as the workspace's (ssplit = 1,
// the drivers' copies, ps = plane_w); the workspace W holds the padded
// fields and the CFL partials (work_fields_bytes).
template <typename T, int N, class PS, class PX, class PZ>
int step_impl(const sles_plan_s* p, int scheme, const sles_step_params* prm, void* uh_, void* uh0_, void* uh1_,
              void* aux_, long ps, int ssplit, void* u_out, void* curl_out, double* scalars, void* work,
              hipStream_t s, void* stages, const DriveArgs* drive, const Batch* batch) {
  using F = Fused<T, N, PS, PX, PZ>;
  const int nm = batch ? batch->count : 1;
  const Members mb = batch ? batch->mb : Members{0, 0, 0, nullptr};
  cpx<T>* uh = (cpx<T>*)uh_;
  cpx<T>* uh0 = (cpx<T>*)uh0_;
  cpx<T>* uh1 = (cpx<T>*)uh1_;
  cpx<T>* aux = (cpx<T>*)aux_;
  cpx<T>* W = (cpx<T>*)work;
  const int les = prm->les_model;
  const int nw = les ? 9 : 6;
  double* part = (double*)(W + nw * cplx_w(p));
  const bool ddt = prm->dt_device != 0 || drive;
  StepArgs a{};
  a.dt = prm->dt;
  a.dt_dev = drive ? drive->state + DRV_DT_APPLY : ddt ? scalars + 1 : nullptr;
  a.nu = prm->nu;
  a.forcing = prm->forcing;
  a.stages = scheme == SLES_SCHEME_RK4 ? stages : nullptr;
  const T coeff = (T)prm->sgs_coeff;
  int nst = 1;
  int epi = EPI_RHS;
  switch (scheme) {
    case SLES_SCHEME_RK4: nst = 4; epi = EPI_RK4; break;
    case SLES_SCHEME_RK3: nst = 3; epi = EPI_RK3; break;
    case SLES_SCHEME_ETD2: nst = 1; epi = EPI_ETD2; break;
    case SLES_SCHEME_RHS: nst = 1; epi = EPI_RHS; break;
    case SLES_SCHEME_NONLINEAR: nst = 1; epi = EPI_NONLIN; break;
    default: return fail("sles_step: unknown scheme");
  }
  static const double rk4a[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  static const double rk4b[4] = {0.5, 0.5, 1.0, 0.0};
  for (int st = 0; st < nst; ++st) {
    a.scheme = epi;
    a.stage = st;
    a.ca = rk4a[st];
    a.cb = rk4b[st];
    double* pp = (ddt && st == 0) ? part : nullptr;
    const bool fuse = F::FUSE_X && p->tune_fuse_x && scheme == SLES_SCHEME_RK4;
    // the x-inverse ran in the previous x-forward (of this step, or of the
    // driver's previous step)
    const bool fused_in = fuse && (st > 0 || (drive && drive->xinv_in));
    // stage 3 runs the driver's next step's x-inverse, or the physical pass's
    const bool next3 = fuse && drive && drive->xinv_out;
    if (F::stage(p, les, coeff, uh, ps, ssplit, W, pp, s, fused_in, mb, nm)) return -1;
    if (pp) {
      // two levels: up to 64 chunks of the per-block maxima, then one block
      // (the partials area holds N^2 + 64 doubles; the z grid is at most N^2)
      ProfScope ps_(PK_CFL, s);
      constexpr int NB = F::ZB < 64 ? F::ZB : 64;
      constexpr int CHUNK = (F::ZB + NB - 1) / NB;
      double* part2 = part + F::ZB;
      hipLaunchKernelGGL(k_nanmax_chunks, dim3(NB, nm), dim3(256), 0, s, (const double*)part, F::ZB, CHUNK, part2,
                         mb);
      if (drive)
        hipLaunchKernelGGL(k_cfl_drive, dim3(1, nm), dim3(64), 0, s, (const double*)part2, NB, prm->cfl,
                           drive->cfl_first, prm->dx, prm->nu, drive->state, mb);
      else
        hipLaunchKernelGGL(k_cfl, dim3(1), dim3(64), 0, s, (const double*)part2, NB, prm->cfl, prm->dx, prm->nu,
                           scalars);
      if (check_launch("cfl")) return -1;
    }
    if (F::finish(p, les, uh, uh0, uh1, aux, ps, ssplit, W, a, s, fuse && (st < 3 || next3),
                  fuse && st == 3 && !next3 && (u_out || curl_out), mb, nm))
      return -1;
  }
  // the last fused RK4 x-forward already ran the physical pass's x-inverse
  // (unless it ran the next step's)
  const bool phys_fused = F::FUSE_X && p->tune_fuse_x && scheme == SLES_SCHEME_RK4 && !(drive && drive->xinv_out);
  if (u_out || curl_out) return F::physical(p, uh, ps, ssplit, W, u_out, curl_out, s, phys_fused);
  return 0;
}

// Line plans selectable per kernel group (plan->tune_*): the default plan or
// the alternative one (LinePlan4: radix-4, or 6 points per thread for the
// sizes with a factor 3) where one exists for N.
template <int N>
struct HasP4 {
  // sizes with an alternative plan (LinePlan4)
  static constexpr bool value = (N == 32 || N == 64 || N == 128 || N == 256 || N == 512 || N == 12 || N == 24 ||
                                 N == 48 || N == 96 || N == 192 || N == 384);
};

// Sizes whose strided passes have narrow small-grid tiles to choose
// (tune_narrow; the tile widths differ from 64^3 to 128^3 only).
template <int N>
struct HasNarrow {
  static constexpr bool value = N == 64 || N == 96 || N == 128;
};

template <typename T, int N>
int step_t(const sles_plan_s* p, int scheme, const sles_step_params* prm, void* uh, void* uh0, void* uh1, void* aux,
           long ps, int ssplit, void* u_out, void* curl_out, double* scalars, void* work, hipStream_t s,
           void* stages, const DriveArgs* drive, const Batch* batch) {
  using D = LinePlan<N>;
  if constexpr (HasP4<N>::value) {
    using Q = LinePlan4<N>;
    const int sel = (p->tune_strided ? 1 : 0) | (p->tune_xfwd ? 2 : 0) | (p->tune_z ? 4 : 0);
#define SLES_SEL(K, A, B, Cz) \
  case K:                     \
    return step_impl<T, N, A, B, Cz>(p, scheme, prm, uh, uh0, uh1, aux, ps, ssplit, u_out, curl_out, scalars, \
                                     work, s, stages, drive, batch);
    if constexpr (HasNarrow<N>::value) {
      if (p->tune_narrow) {
        // the strided and x-forward passes in narrow tiles (the z pass has no tiles)
        using ND = Narrow<D>;
        using NQ = Narrow<Q>;
        switch (sel) {
          SLES_SEL(0, ND, ND, D)
          SLES_SEL(1, NQ, ND, D)
          SLES_SEL(2, ND, NQ, D)
          SLES_SEL(3, NQ, NQ, D)
          SLES_SEL(4, ND, ND, Q)
          SLES_SEL(5, NQ, ND, Q)
          SLES_SEL(6, ND, NQ, Q)
          SLES_SEL(7, NQ, NQ, Q)
        }
      }
    }
    switch (sel) {
      SLES_SEL(0, D, D, D)
      SLES_SEL(1, Q, D, D)
      SLES_SEL(2, D, Q, D)
      SLES_SEL(3, Q, Q, D)
      SLES_SEL(4, D, D, Q)
      SLES_SEL(5, Q, D, Q)
      SLES_SEL(6, D, Q, Q)
      SLES_SEL(7, Q, Q, Q)
    }
#undef SLES_SEL
  }
  return step_impl<T, N, D, D, D>(p, scheme, prm, uh, uh0, uh1, aux, ps, ssplit, u_out, curl_out, scalars, work,
                                  s, stages, drive, batch);
}

// Per-precision instantiation (sles_rhs_f64.hip / sles_rhs_f32.hip).
// ps, ssplit: kx-plane stride (elements) and arrangement (0: the
// reference's, 1: the workspace's) of the RK-state arrays uh, uh0, uh1, aux,
// stages.
int step_f64(const sles_plan_s* p, int scheme, const sles_step_params* prm, void* uh, void* uh0, void* uh1,
             void* aux, long ps, int ssplit, void* u_out, void* curl_out, double* scalars, void* work,
             hipStream_t s, void* stages, const DriveArgs* drive = nullptr, const Batch* batch = nullptr);
int step_f32(const sles_plan_s* p, int scheme, const sles_step_params* prm, void* uh, void* uh0, void* uh1,
             void* aux, long ps, int ssplit, void* u_out, void* curl_out, double* scalars, void* work,
             hipStream_t s, void* stages, const DriveArgs* drive = nullptr, const Batch* batch = nullptr);

}  // namespace sles

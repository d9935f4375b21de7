// Hand-written adjoint of the fused RHS: out = J(Y)^T mu, where J is the
// Jacobian of computeRHS (src/RK4_method.py:38-71 with the Smagorinsky /
// Vreman closures, src/smagorinsky.py:24-39, src/vreman.py:24-123) with
// respect to U_hat at the state Y, and the FP64 sum of nu_bar * d nu/d Cs
// (the Cs gradient), for an adjoint mu of the RHS.
//
// Convention: complex arrays are real vectors (Re, Im interleaved); the
// adjoint of a real-linear map is its transpose in that inner product, which
// is the layout torch uses for gradients of complex tensors.  Transposes used:
//   c2r (scaled 1/N^3, numpy irfftn semantics):  C2R^T g = (Wz/N^3) R2C(g)
//   r2c (unscaled):                               R2C^T G = C2R_unscaled(G / Wz)
//   multiplication by i*K:                        multiplication by -i*K
//   projection Pi = I - K K^T/K^2 (per k):        Pi (symmetric)
// with Wz = 2 for 0 < kz < N/2 and 1 on the kz = 0, N/2 planes.
//
// Passes (workspace A of 17 spectral fields, plus A17-19 for a combined mu):
//   1 x-inverse   Y -> A0-2 = Y, A3-5 = i Kx Y ;  g = dealias Pi mu -> A9-11 = g,
//                 A12-14 = -2i Kx g ; mu = c1 x + c2 z formed here -> A17-19
//   2 y-inverse   A0-8 as the forward LES pass; A9-11 = g, A12 = tau_bar_xx,
//                 A13 = tau_bar_xy, A14 = tau_bar_xz (Kx part), A15 = tau_bar_yy,
//                 A16 = tau_bar_yz (Ky part)
//   3 z kernel    c2r of u, grad u (scaled) and of n_bar, tau_bar (unscaled,
//                 / Wz, Kz parts added); pointwise adjoint of
//                 omega / u x omega / nu_R / tau; r2c of u_bar, a_bar;
//                 A0-2 = R2C(u_bar_j) - i Kz R2C(a_bar_zj), A3-5 = a_bar_xj,
//                 A6-8 = a_bar_yj; Cs partial sums
//   4 y-forward   A0-2 += -i Ky F_y(A6-8); A3-5 in place
//   5 x-forward   out_j (=|+=) (Wz/N^3) (F_x(A_j) - i Kx F_x(A_{3+j}))
//                 [+ (f [k != 0] - nu K^2) mu_j]
#pragma once

#include <type_traits>

#include "sles_rhs.h"

namespace sles {
constexpr int kAdjFields = 20;  // spectral fields of the adjoint workspace
}

namespace sles {

template <typename T>
__device__ __forceinline__ cpx<T> mul_mis(T s, cpx<T> a) {  // (-i s) a
  return {s * a.y, -s * a.x};
}

// Where the x-inverse and x-forward passes take the adjoint mu from:
//   etd2 = 0: mu = c1 lam1 + c2 lam2 (lam2 may be null; c1 = 1, lam2 = null
//             is mu = lam1 as given);
//   etd2 = 1: mu = w1 lam1 - w0 lam2 formed per mode (ETD2 backward step),
//             with the output out = E lam1 + J^T mu (E = e^{hL}, L = -nu K^2,
//             w1 = h (phi1 + phi2), w0 = h phi2).
struct AdjSrc {
  const void* lam1;
  const void* lam2;  // may be null (treated as zero)
  double c1, c2;
  double h;
  int etd2;
};
// mw (etd2 = 0 only, may be null): where the x-inverse pass stores the mu it
// forms, for the x-forward pass's linear term (which then reads one array).
template <typename T>
struct MuSrc {
  const cpx<T>* l1;
  const cpx<T>* l2;
  T c1, c2;
  double h, nu;
  cpx<T>* mw;
  __device__ __forceinline__ void weights(long k2, T& w1, T& w0, T& E) const {
    const double z = -h * nu * (double)k2;
    const double p1 = phi1_d(z), p2 = phi2_d(z);
    w1 = (T)(h * (p1 + p2));
    w0 = (T)(h * p2);
    E = (T)exp(z);
  }
  // mu of component j at element e (offset within the field); ETD2: weights w1, w0 of that mode
  template <bool ETD2>
  __device__ __forceinline__ cpx<T> at(int j, long C, long e, T w1, T w0) const {
    if constexpr (!ETD2) {
      cpx<T> r = c1 * l1[j * C + e];
      if (l2) r = r + c2 * l2[j * C + e];
      return r;
    } else {
      cpx<T> r = w1 * l1[j * C + e];
      if (l2) r = r - w0 * l2[j * C + e];
      return r;
    }
  }
};

// out[b] = sum of part[b*chunk, (b+1)*chunk)  (first level, fixed order)
static __global__ void k_sum_chunks(const double* part, int np, int chunk, double* out) {
  __shared__ double sh[256];
  double v = 0.0;
  const int lo = blockIdx.x * chunk, hi = lo + chunk < np ? lo + chunk : np;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) v += part[i];
  sh[threadIdx.x] = v;
  block_sync();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    block_sync();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// acc[0] += sum of part[0..np)  (one block, fixed order)
static __global__ void k_sum_add(const double* part, int np, double* acc) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  sh[threadIdx.x] = v;
  block_sync();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    block_sync();
  }
  if (threadIdx.x == 0) acc[0] += sh[0];
}

// ---------------------------------------------------------------------------
// 1. x-inverse
// ---------------------------------------------------------------------------
template <typename T, int N, class PL, bool ETD2>
__global__ void __launch_bounds__((StridedShape<T, N, PL>::THREADS))
    k_adj_xinv(const cpx<T>* __restrict__ Y, MuSrc<T> mu, cpx<T>* __restrict__ A, long C,
               const cpx<T>* __restrict__ tw, int xcd) {
  using S = StridedShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, +1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  // FP64 radix-8 ETD2 variant: table twiddles (the ETD2 weights need the registers)
  const auto twr = TwPick<T, PL, !(ETD2 && sizeof(T) == 8 && PL::P >= 8)>::make(tw, t);
  const LdsColumn<TK> map{col.c};
  const int ky = kfreq((int)(col.g / NH), N), kz = (int)(col.g % NH);
  auto at = [&](int m) { return (long)(t + m * NT) * Cols<N, TK>::M + col.g; };
  cpx<T> v[P], w[P];
  // forward fields (as the forward LES x pass)
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
      v[m] = Y[j * C + at(m)];
      w[m] = mul_i((T)kfreq(t + m * NT, N) * v[m]);
    }
    FFT::run(v, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) A[j * C + at(m)] = v[m];
    FFT::run(w, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) A[(3 + j) * C + at(m)] = w[m];
  }
  // g = dealias * Pi mu; s = (K . mu) / K^2.  With room in the register file
  // (FP32, or 4 points per thread) the three mu components stay in registers
  // for the second loop; otherwise they are read again (with the ETD2 weights
  // of each mode kept).
  constexpr bool HOLD = sizeof(T) == 4 || P <= 4;
  cpx<T> s[P];
  cpx<T> mh0[HOLD ? P : 1], mh1[HOLD ? P : 1], mh2[HOLD ? P : 1];
  T w1[HOLD ? 1 : P], w0[HOLD ? 1 : P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int kx = kfreq(t + m * NT, N);
    const long k2 = (long)kx * kx + (long)ky * ky + (long)kz * kz;
    const T inv = T(1) / (T)(k2 == 0 ? 1 : k2);
    const long e = at(m);
    T a1 = T(0), a0 = T(0), E;
    if constexpr (ETD2) mu.weights(k2, a1, a0, E);
    const cpx<T> m0 = mu.template at<ETD2>(0, C, e, a1, a0), m1 = mu.template at<ETD2>(1, C, e, a1, a0),
                 m2 = mu.template at<ETD2>(2, C, e, a1, a0);
    if constexpr (!ETD2) {
      if (mu.mw && col.valid) {
        mu.mw[e] = m0;
        mu.mw[C + e] = m1;
        mu.mw[2 * C + e] = m2;
      }
    }
    if constexpr (HOLD) {
      mh0[m] = m0;
      mh1[m] = m1;
      mh2[m] = m2;
    } else {
      w1[m] = a1;
      w0[m] = a0;
    }
    s[m] = inv * (((T)kx * m0 + (T)ky * m1) + (T)kz * m2);
  }
  // component j of g and of -2i Kx g, transformed and stored
  auto emit = [&](int j, const cpx<T>(&mj)[P]) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int kx = kfreq(t + m * NT, N);
      const KIndex k = {kx, ky, kz};
      const T keep = dealias_keep(k, N) ? T(1) : T(0);
      const T kj = (T)(j == 0 ? kx : (j == 1 ? ky : kz));
      v[m] = keep * (mj[m] - kj * s[m]);
      w[m] = mul_mis(T(2) * (T)kx, v[m]);
    }
    FFT::run(v, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) A[(9 + j) * C + at(m)] = v[m];
    FFT::run(w, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) A[(12 + j) * C + at(m)] = w[m];
  };
  if constexpr (HOLD) {
    // straight-line: each held component dies after its use
    emit(0, mh0);
    emit(1, mh1);
    emit(2, mh2);
  } else {
#pragma unroll 1
    for (int j = 0; j < 3; ++j) {
      cpx<T> mj[P];
#pragma unroll
      for (int m = 0; m < P; ++m) mj[m] = mu.template at<ETD2>(j, C, at(m), w1[m], w0[m]);
      emit(j, mj);
    }
  }
}

// ---------------------------------------------------------------------------
// 2. y-inverse (in place)
// ---------------------------------------------------------------------------
template <typename T, int N, class PL>
__global__ void __launch_bounds__((StridedShape<T, N, PL>::THREADS))
    k_adj_yinv(cpx<T>* A, long C, const cpx<T>* __restrict__ tw, int xcd) {
  using S = StridedShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, +1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const long ix = col.g / NH, iz = col.g % NH;
  auto at = [&](int m) { return ix * (long)N * NH + (long)(t + m * NT) * NH + iz; };
  auto ky = [&](int m) { return (T)kfreq(t + m * NT, N); };
  cpx<T> v[P], w[P];
  auto load = [&](int f, cpx<T>(&x)[P]) {
#pragma unroll
    for (int m = 0; m < P; ++m) x[m] = A[f * C + at(m)];
  };
  auto store = [&](int f, const cpx<T>(&x)[P]) {
    if (col.valid)
#pragma unroll
      for (int m = 0; m < P; ++m) A[f * C + at(m)] = x[m];
  };
  // forward fields: A0-2 -> A0-2 and A6-8 = i Ky (.), A3-5 in place
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
    load(j, v);
#pragma unroll
    for (int m = 0; m < P; ++m) w[m] = mul_i(ky(m) * v[m]);
    FFT::run(v, lds, map, t, twr);
    store(j, v);
    FFT::run(w, lds, map, t, twr);
    store(6 + j, w);
  }
#pragma unroll 1
  for (int j = 3; j < 6; ++j) {
    load(j, v);
    FFT::run(v, lds, map, t, twr);
    store(j, v);
  }
  // adjoint: g_x (A9) -> A9, and tau_bar_xy = A13 - 2i Ky g_x
  load(9, v);
#pragma unroll
  for (int m = 0; m < P; ++m) w[m] = mul_mis(T(2) * ky(m), v[m]);
  FFT::run(v, lds, map, t, twr);
  store(9, v);
  load(13, v);
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = v[m] + w[m];
  FFT::run(v, lds, map, t, twr);
  store(13, v);
  // g_y (A10) -> A10, tau_bar_yy = -2i Ky g_y -> A15 ; g_z (A11) -> A11, -2i Ky g_z -> A16
#pragma unroll 1
  for (int j = 0; j < 2; ++j) {
    load(10 + j, v);
#pragma unroll
    for (int m = 0; m < P; ++m) w[m] = mul_mis(T(2) * ky(m), v[m]);
    FFT::run(v, lds, map, t, twr);
    store(10 + j, v);
    FFT::run(w, lds, map, t, twr);
    store(15 + j, w);
  }
  // tau_bar_xx (A12), tau_bar_xz Kx part (A14) in place
#pragma unroll 1
  for (int f = 12; f <= 14; f += 2) {
    load(f, v);
    FFT::run(v, lds, map, t, twr);
    store(f, v);
  }
}

// ---------------------------------------------------------------------------
// 3. z kernel
// ---------------------------------------------------------------------------
// Z_k = a_k + i b_k of two half-spectrum operands, Hermitian completion as
// c2r: fa(k), fb(k) give the operands at bins k <= N/2 (each evaluated once,
// from memory); bins k > N/2 take conj of bin N - k from the line's LDS row
// (pmap: unpermuted layout; free on entry and on return).  HAS_B false: b = 0.
template <typename T, int N, int P, int NT, bool WAVE, class Map, class FA, class FB>
__device__ __forceinline__ void complete_pair(cpx<T> (&v)[P], const FA& fa, const FB& fb, bool has_b, int t,
                                              cpx<T>* lds, const Map& pmap) {
  // both operands' loads are issued before the first LDS round
  cpx<T> ha[P / 2 + 1], hb[P / 2 + 1];
#pragma unroll
  for (int m = 0; m <= P / 2; ++m) {
    const int k = t + m * NT;
    if (k <= N / 2) {
      ha[m] = fa(k);
      if (has_b) hb[m] = fb(k);
    }
  }
  auto round = [&](const cpx<T>(&h)[P / 2 + 1], bool second) {
#pragma unroll
    for (int m = 0; m <= P / 2; ++m)
      if (t + m * NT <= N / 2) lds[pmap(t + m * NT)] = h[m];
    line_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int k = t + m * NT;
      cpx<T> x = lds[pmap(k <= N / 2 ? k : N - k)];
      if (k > N / 2)
        x = conj(x);
      else if (k == 0 || k == N / 2)
        x.y = T(0);
      v[m] = second ? cpx<T>{v[m].x - x.y, v[m].y + x.x} : x;
    }
    line_sync<WAVE>();
  };
  round(ha, false);
  if (has_b) round(hb, true);
}

// A pair of half-spectra rows A, B (row pointers in bounds even for an
// invalid line, never stored), optionally multiplied by i*kz first.
template <typename T, int N, int P, int NT, bool WAVE, class Map>
__device__ __forceinline__ void load_pair_rows(cpx<T> (&v)[P], const cpx<T>* A, bool dzA, const cpx<T>* B, bool dzB,
                                               int t, cpx<T>* lds, const Map& pmap) {
  auto fa = [&](int k) { return dzA ? mul_i((T)k * A[k]) : A[k]; };
  auto fb = [&](int k) { return dzB ? mul_i((T)k * B[k]) : B[k]; };
  complete_pair<T, N, P, NT, WAVE>(v, fa, fb, true, t, lds, pmap);
}

// One c2r operand at bin k <= N/2: (F1 [+ (-2i kz) F2]) / Wz(k).
template <typename T, int N>
__device__ __forceinline__ cpx<T> adj_term(const cpx<T>* F1, const cpx<T>* F2, int k) {
  cpx<T> a = F1 ? F1[k] : cpx<T>{T(0), T(0)};
  if (F2) a = a + mul_mis(T(2) * (T)k, F2[k]);
  return (k == 0 || k == N / 2) ? a : T(0.5) * a;
}

template <typename T, int N, int P, int NT, bool WAVE, class Map>
__device__ __forceinline__ void load_adj_pair(cpx<T> (&v)[P], const cpx<T>* A1, const cpx<T>* A2, const cpx<T>* B1,
                                              const cpx<T>* B2, int t, cpx<T>* lds, const Map& pmap) {
  auto fa = [&](int k) { return adj_term<T, N>(A1, A2, k); };
  auto fb = [&](int k) { return adj_term<T, N>(B1, B2, k); };
  complete_pair<T, N, P, NT, WAVE>(v, fa, fb, B1 || B2, t, lds, pmap);
}

__device__ __forceinline__ void block_sum_store(double v, double* out) {
  __shared__ double s[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) s[wid] = v;
  block_sync();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) r += s[i];
    out[blockIdx.x] = r;
  }
}

template <typename T, int N, class PL, int MODEL>
__global__ void __launch_bounds__(ZLines<PL>::THREADS)
    k_adj_z(cpx<T>* A, long C, T scale, T coeff, T dcoeff, double* cs_part, const cpx<T>* __restrict__ tw) {
  constexpr int P = PL::P, NT = PL::NT, NH = N / 2 + 1, KH = P / 2 + 1;
  constexpr int LPB = ZLines<PL>::LPB;
  constexpr bool WV = WaveLocal<PL>::value;  // each line inside one wavefront: no block barriers
  using IFFT = LineFFTP<T, PL, +1, WV, ZPPB<T, N, PL>>;
  using FFT = LineFFTP<T, PL, -1, WV, ZPPB<T, N, PL>>;
  __shared__ cpx<T> lds[LPB * LdsLine<T, N>::STRIDE + ZPPB<T, N, PL>];
  const int line = threadIdx.x / NT, t = threadIdx.x % NT;
  const long row = (long)blockIdx.x * LPB + line;
  const bool valid = row < (long)N * N;
  const long base = (valid ? row : 0) * NH;
  const LdsLine<T, N> map{line * LdsLine<T, N>::STRIDE};
  const TwRegs<T, PL> twr(tw, t);  // read once, used by all 17 transforms
  cpx<T>* w[17];
#pragma unroll
  for (int i = 0; i < 17; ++i) w[i] = A + i * C + base;

  T u0[P], u1[P], u2[P], a[3][3][P];
  cpx<T> v[P];
#define SLES_INV(PA, DA, PB, DB, RA, RB)                           \
  load_pair_rows<T, N, P, NT, WV>(v, PA, DA, PB, DB, t, lds, map.plain()); \
  IFFT::run(v, lds, map, t, twr);                                   \
  _Pragma("unroll") for (int m = 0; m < P; ++m) {                  \
    RA[m] = scale * v[m].x;                                        \
    RB[m] = scale * v[m].y;                                        \
  }
  SLES_INV(w[0], false, w[1], false, u0, u1)
  SLES_INV(w[2], false, w[0], true, u2, a[2][0])
  SLES_INV(w[1], true, w[2], true, a[2][1], a[2][2])
  SLES_INV(w[3], false, w[4], false, a[0][0], a[0][1])
  SLES_INV(w[5], false, w[6], false, a[0][2], a[1][0])
  SLES_INV(w[7], false, w[8], false, a[1][1], a[1][2])
#undef SLES_INV
  // adjoint sources (unscaled c2r of (.)/Wz)
  T nb0[P], nb1[P], nb2[P], t11[P], t12[P], t22[P], t13[P], t23[P], t33[P];
#define SLES_AINV(A1, A2, B1, B2, RA, RB)                          \
  load_adj_pair<T, N, P, NT, WV>(v, A1, A2, B1, B2, t, lds, map.plain()); \
  IFFT::run(v, lds, map, t, twr);                                   \
  _Pragma("unroll") for (int m = 0; m < P; ++m) {                  \
    RA[m] = v[m].x;                                                \
    RB[m] = v[m].y;                                                \
  }
  SLES_AINV(w[9], nullptr, w[10], nullptr, nb0, nb1)
  SLES_AINV(w[11], nullptr, w[12], nullptr, nb2, t11)
  SLES_AINV(w[13], nullptr, w[15], nullptr, t12, t22)
  SLES_AINV(w[14], w[9], w[16], w[10], t13, t23)
  T zq[P];
  SLES_AINV((const cpx<T>*)nullptr, w[11], (const cpx<T>*)nullptr, (const cpx<T>*)nullptr, t33, zq)
#undef SLES_AINV

  // pointwise adjoint
  double cs_acc = 0.0;
  T ub0[P], ub1[P], ub2[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    T g[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) g[i][j] = a[i][j][m];
    const T w0 = g[1][2] - g[2][1], w1 = g[2][0] - g[0][2], w2 = g[0][1] - g[1][0];
    // n = u x omega: u_bar = omega x n_bar, omega_bar = n_bar x u
    ub0[m] = w1 * nb2[m] - w2 * nb1[m];
    ub1[m] = w2 * nb0[m] - w0 * nb2[m];
    ub2[m] = w0 * nb1[m] - w1 * nb0[m];
    const T wb0 = nb1[m] * u2[m] - nb2[m] * u1[m];
    const T wb1 = nb2[m] * u0[m] - nb0[m] * u2[m];
    const T wb2 = nb0[m] * u1[m] - nb1[m] * u0[m];
    T gb[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gb[i][j] = T(0);
    gb[1][2] += wb0;
    gb[2][1] -= wb0;
    gb[2][0] += wb1;
    gb[0][2] -= wb1;
    gb[0][1] += wb2;
    gb[1][0] -= wb2;
    if constexpr (MODEL != SLES_LES_NONE) {
      // shape q = nu / coeff and d q / d a
      T q, dq[3][3];
      if constexpr (MODEL == SLES_LES_SMAGORINSKY) {
        const T s12 = g[0][1] + g[1][0], s13 = g[0][2] + g[2][0], s23 = g[1][2] + g[2][1];
        const T ss = g[0][0] * g[0][0] + T(0.5) * (s12 * s12) + g[1][1] * g[1][1] + T(0.5) * (s13 * s13) +
                     T(0.5) * (s23 * s23) + g[2][2] * g[2][2];
        q = sqrt(T(2) * ss);
        const T iq = q > T(0) ? T(1) / q : T(0);
        dq[0][0] = T(2) * g[0][0] * iq;
        dq[1][1] = T(2) * g[1][1] * iq;
        dq[2][2] = T(2) * g[2][2] * iq;
        dq[0][1] = dq[1][0] = s12 * iq;
        dq[0][2] = dq[2][0] = s13 * iq;
        dq[1][2] = dq[2][1] = s23 * iq;
      } else {
        T b[3][3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int r = 0; r < 3; ++r) b[p][r] = g[0][p] * g[0][r] + g[1][p] * g[1][r] + g[2][p] * g[2][r];
        const T B = b[0][0] * b[1][1] - b[0][1] * b[0][1] + b[0][0] * b[2][2] - b[0][2] * b[0][2] +
                    b[1][1] * b[2][2] - b[1][2] * b[1][2];
        const T sb = sqrt(B);
        const T aa = b[0][0] + b[1][1] + b[2][2];
        const bool on = (aa > T(1e-4) && sb > aa / T(1e7));
        // as the forward (src/vreman.py:119-123): B < 0 (round-off on a
        // near-rank-1 gradient) gives sqrt -> NaN and NaN * 0 stays NaN
        q = sb * (on ? sqrt(T(1) / aa) : T(0));
        // d q / d a_mr = q ((a G)_mr / B - a_mr / aa),  G = tr(beta) I - beta
        const T iB = (on && B > T(0)) ? T(1) / B : T(0), ia = on ? T(1) / aa : T(0);
#pragma unroll
        for (int mm = 0; mm < 3; ++mm)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            T ag = T(0);
#pragma unroll
            for (int c = 0; c < 3; ++c) ag += g[mm][c] * ((c == r ? aa : T(0)) - b[c][r]);
            dq[mm][r] = q * (ag * iB - g[mm][r] * ia);
          }
      }
      const T nu = coeff * q;
      const T tb[3][3] = {{t11[m], t12[m], t13[m]}, {t12[m], t22[m], t23[m]}, {t13[m], t23[m], t33[m]}};
      // tau_ii = a_ii nu ; tau_ij = (a_ij + a_ji)/2 nu (one field for i < j)
      T nub = T(0);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        gb[i][i] += nu * tb[i][i];
        nub += g[i][i] * tb[i][i];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) {
          const T h = T(0.5) * nu * tb[i][j];
          gb[i][j] += h;
          gb[j][i] += h;
          nub += T(0.5) * (g[i][j] + g[j][i]) * tb[i][j];
        }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gb[i][j] += nub * coeff * dq[i][j];
      cs_acc += (double)nub * (double)dcoeff * (double)q;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) a[i][j][m] = gb[i][j];
  }
  if (cs_part) block_sum_store(valid ? cs_acc : 0.0, cs_part);

  // forward r2c of u_bar (3), a_bar (9)
  cpx<T> Ah[KH], Bh[KH], s0[KH], s1[KH], s2[KH];
  cpx<T>* o[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) o[i] = A + i * C + base;
  auto kok = [&](int m) { return t + m * NT <= N / 2; };
#define SLES_FWD(QA, QB)                                               \
  _Pragma("unroll") for (int m = 0; m < P; ++m) v[m] = {QA[m], QB[m]}; \
  FFT::run(v, lds, map, t, twr);                                        \
  split_pair<T, N, P, NT, WV>(v, lds, map.plain(), t, Ah, Bh);
  SLES_FWD(a[2][0], a[2][1])
#pragma unroll
  for (int m = 0; m < KH; ++m) {
    s0[m] = Ah[m];
    s1[m] = Bh[m];
  }
  SLES_FWD(a[2][2], ub0)
#pragma unroll
  for (int m = 0; m < KH; ++m) {
    s2[m] = Ah[m];
    if (valid && kok(m)) o[0][t + m * NT] = Bh[m] + mul_mis((T)(t + m * NT), s0[m]);
  }
  SLES_FWD(ub1, ub2)
#pragma unroll
  for (int m = 0; m < KH; ++m)
    if (valid && kok(m)) {
      o[1][t + m * NT] = Ah[m] + mul_mis((T)(t + m * NT), s1[m]);
      o[2][t + m * NT] = Bh[m] + mul_mis((T)(t + m * NT), s2[m]);
    }
  SLES_FWD(a[0][0], a[0][1])
#pragma unroll
  for (int m = 0; m < KH; ++m)
    if (valid && kok(m)) {
      o[3][t + m * NT] = Ah[m];
      o[4][t + m * NT] = Bh[m];
    }
  SLES_FWD(a[0][2], a[1][0])
#pragma unroll
  for (int m = 0; m < KH; ++m)
    if (valid && kok(m)) {
      o[5][t + m * NT] = Ah[m];
      o[6][t + m * NT] = Bh[m];
    }
  SLES_FWD(a[1][1], a[1][2])
#pragma unroll
  for (int m = 0; m < KH; ++m)
    if (valid && kok(m)) {
      o[7][t + m * NT] = Ah[m];
      o[8][t + m * NT] = Bh[m];
    }
#undef SLES_FWD
}

// ---------------------------------------------------------------------------
// 4. y-forward (in place): A0-2 += -i Ky F_y(A6-8); A3-5 transformed
// ---------------------------------------------------------------------------
template <typename T, int N, class PL>
__global__ void __launch_bounds__((StridedShape<T, N, PL>::THREADS))
    k_adj_yfwd(cpx<T>* A, long C, const cpx<T>* __restrict__ tw, int xcd) {
  using S = StridedShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, -1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  const TwRegs<T, PL> twr(tw, t);  // read once per block, used by every transform
  const LdsColumn<TK> map{col.c};
  const long ix = col.g / NH, iz = col.g % NH;
  auto at = [&](int m) { return ix * (long)N * NH + (long)(t + m * NT) * NH + iz; };
  cpx<T> v[P], acc[P];
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = A[(6 + j) * C + at(m)];
    FFT::run(v, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) acc[m] = mul_mis((T)kfreq(t + m * NT, N), v[m]);
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = A[j * C + at(m)];
    FFT::run(v, lds, map, t, twr);
    if (col.valid)
#pragma unroll
      for (int m = 0; m < P; ++m) A[j * C + at(m)] = v[m] + acc[m];
  }
#pragma unroll 1
  for (int j = 3; j < 6; ++j) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = A[j * C + at(m)];
    FFT::run(v, lds, map, t, twr);
    if (col.valid)
#pragma unroll
      for (int m = 0; m < P; ++m) A[j * C + at(m)] = v[m];
  }
}

// ---------------------------------------------------------------------------
// 5. x-forward + output
// ---------------------------------------------------------------------------
struct AdjArgs {
  double nu, forcing;
  int linear;      // 1: add (forcing [k != 0] - nu K^2) mu; 2: add forcing [k != 0] mu; 0: none
  int accumulate;  // out += result instead of out = result
};

template <typename T, int N, class PL, bool ETD2>
__global__ void __launch_bounds__((StridedShape<T, N, PL>::THREADS))
    k_adj_xfwd(const cpx<T>* __restrict__ A, long C, MuSrc<T> mu, cpx<T>* out, AdjArgs a,
               const cpx<T>* __restrict__ tw, int xcd) {
  using S = StridedShape<T, N, PL>;
  constexpr int P = S::P, NT = S::NT, TK = S::TK, NH = N / 2 + 1;
  constexpr int PPB = N * TK;  // ping-pong exchange buffers
  using FFT = LineFFTP<T, PL, -1, false, PPB>;
  __shared__ cpx<T> lds[N * TK + PPB];
  const Cols<N, TK> col(xcd);
  const int t = threadIdx.x / TK;
  // FP64 radix-8 ETD2 variant: table twiddles (the ETD2 weights need the registers)
  const auto twr = TwPick<T, PL, !(ETD2 && sizeof(T) == 8 && PL::P >= 8)>::make(tw, t);
  const LdsColumn<TK> map{col.c};
  auto at = [&](int m) { return (long)(t + m * NT) * Cols<N, TK>::M + col.g; };
  const int ky = kfreq((int)(col.g / NH), N), kz = (int)(col.g % NH);
  const T wz = (kz == 0 || kz == N / 2) ? T(1) : T(2);
  const T sc = wz * T(1.0 / ((double)N * N * N));
  cpx<T> r[P], v[P];
#pragma unroll 1
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = A[(3 + j) * C + at(m)];
    FFT::run(v, lds, map, t, twr);
#pragma unroll
    for (int m = 0; m < P; ++m) r[m] = mul_mis((T)kfreq(t + m * NT, N), v[m]);
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = A[j * C + at(m)];
    FFT::run(v, lds, map, t, twr);
    if (col.valid) {
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const long e = j * C + at(m);
        const int kx = kfreq(t + m * NT, N);
        const long k2 = (long)kx * kx + (long)ky * ky + (long)kz * kz;
        cpx<T> res = sc * (v[m] + r[m]);
        if constexpr (ETD2) {
          // e^{hL} lam_{n+1} + J^T mu, with the forcing part of the linear term
          T w1, w0, E;
          mu.weights(k2, w1, w0, E);
          if (a.forcing > 0.0 && k2 != 0) res = res + (T)a.forcing * mu.template at<true>(j, C, at(m), w1, w0);
          out[e] = E * mu.l1[e] + res;
        } else {
          if (a.linear) {
            const double visc = a.linear == 1 ? a.nu * (double)k2 : 0.0;
            const T phi = (T)((a.forcing > 0.0 && k2 != 0 ? a.forcing : 0.0) - visc);
            res = res + phi * mu.l1[e];  // mu as given, or as stored by the x-inverse pass
          }
          out[e] = a.accumulate ? out[e] + res : res;
        }
      }
    }
  }
}

template <typename T, int N, class PS, class PZ>
int adjoint_impl(const sles_plan_s* p, const sles_step_params* prm, const void* y, const AdjSrc& src, void* out,
                 int accumulate, int linear, double dcoeff, double* cs_grad, void* work, hipStream_t s) {
  using SS = StridedShape<T, N, PS>;
  constexpr int ZB = (int)(((long)N * N + ZLines<PZ>::LPB - 1) / ZLines<PZ>::LPB);
  const long C = ncplx(p);
  const cpx<T>* tw = (const cpx<T>*)p->tw;
  const int xcd = p->tune_xcd;
  cpx<T>* A = (cpx<T>*)work;
  double* part = (double*)(A + kAdjFields * C);
  constexpr long NCOL = (long)N * (N / 2 + 1);  // columns of every strided pass
  const dim3 gs((unsigned)((NCOL + SS::TK - 1) / SS::TK)), bs(SS::THREADS);
  const T scale = T(1.0 / ((double)N * N * N));
  // a combined mu that the x-forward pass needs is stored once, in A[17..19]
  cpx<T>* const mw = (!src.etd2 && linear && (src.lam2 || src.c1 != 1.0)) ? A + 17 * C : nullptr;
  const MuSrc<T> ms{(const cpx<T>*)src.lam1, (const cpx<T>*)src.lam2, (T)src.c1, (T)src.c2, src.h, prm->nu, mw};
  const MuSrc<T> msx = mw ? MuSrc<T>{mw, nullptr, T(1), T(0), src.h, prm->nu, nullptr} : ms;
  if (src.etd2)
    hipLaunchKernelGGL((k_adj_xinv<T, N, PS, true>), gs, bs, 0, s, (const cpx<T>*)y, ms, A, C, tw, xcd);
  else
    hipLaunchKernelGGL((k_adj_xinv<T, N, PS, false>), gs, bs, 0, s, (const cpx<T>*)y, ms, A, C, tw, xcd);
  hipLaunchKernelGGL((k_adj_yinv<T, N, PS>), gs, bs, 0, s, A, C, tw, xcd);
  double* cp = cs_grad ? part : nullptr;
  switch (prm->les_model) {
    case SLES_LES_SMAGORINSKY:
      hipLaunchKernelGGL((k_adj_z<T, N, PZ, SLES_LES_SMAGORINSKY>), dim3(ZB), dim3(ZLines<PZ>::THREADS), 0, s, A, C,
                         scale, (T)prm->sgs_coeff, (T)dcoeff, cp, tw);
      break;
    case SLES_LES_VREMAN:
      hipLaunchKernelGGL((k_adj_z<T, N, PZ, SLES_LES_VREMAN>), dim3(ZB), dim3(ZLines<PZ>::THREADS), 0, s, A, C,
                         scale, (T)prm->sgs_coeff, (T)dcoeff, cp, tw);
      break;
    default:
      hipLaunchKernelGGL((k_adj_z<T, N, PZ, SLES_LES_NONE>), dim3(ZB), dim3(ZLines<PZ>::THREADS), 0, s, A, C, scale,
                         T(0), T(0), (double*)nullptr, tw);
      cp = nullptr;
  }
  if (cp) {
    // two levels (up to 64 chunks, then one block); the partials area holds
    // N^2 + 64 doubles and the z grid has at most N^2 blocks
    constexpr int NB = ZB < 64 ? ZB : 64;
    constexpr int CHUNK = (ZB + NB - 1) / NB;
    hipLaunchKernelGGL(k_sum_chunks, dim3(NB), dim3(256), 0, s, (const double*)part, ZB, CHUNK, part + ZB);
    hipLaunchKernelGGL(k_sum_add, dim3(1), dim3(64), 0, s, (const double*)(part + ZB), NB, cs_grad);
  }
  hipLaunchKernelGGL((k_adj_yfwd<T, N, PS>), gs, bs, 0, s, A, C, tw, xcd);
  AdjArgs aa{prm->nu, prm->forcing, linear, accumulate};
  if (src.etd2)
    hipLaunchKernelGGL((k_adj_xfwd<T, N, PS, true>), gs, bs, 0, s, (const cpx<T>*)A, C, ms, (cpx<T>*)out, aa, tw, xcd);
  else
    hipLaunchKernelGGL((k_adj_xfwd<T, N, PS, false>), gs, bs, 0, s, (const cpx<T>*)A, C, msx, (cpx<T>*)out, aa, tw,
                       xcd);
  return check_launch("sles_adjoint_rhs");
}

template <typename T, int N>
int adjoint_t(const sles_plan_s* p, const sles_step_params* prm, const void* y, const AdjSrc& mu, void* out,
              int accumulate, int linear, double dcoeff, double* cs_grad, void* work, hipStream_t s) {
  using D = LinePlan<N>;
  if constexpr (HasP4<N>::value) {
    using Q = LinePlan4<N>;
    // the adjoint z kernel holds 21 physical lines per thread: the
    // alternative plan (radix-4, or 6 points for sizes with a factor 3; half
    // the registers of the default plan) in both precisions
    if (p->tune_strided) return adjoint_impl<T, N, Q, Q>(p, prm, y, mu, out, accumulate, linear, dcoeff, cs_grad, work, s);
    return adjoint_impl<T, N, D, Q>(p, prm, y, mu, out, accumulate, linear, dcoeff, cs_grad, work, s);
  }
  return adjoint_impl<T, N, D, D>(p, prm, y, mu, out, accumulate, linear, dcoeff, cs_grad, work, s);
}

int adjoint_f64(const sles_plan_s* p, const sles_step_params* prm, const void* y, const AdjSrc& mu, void* out,
                int accumulate, int linear, double dcoeff, double* cs_grad, void* work, hipStream_t s);
int adjoint_f32(const sles_plan_s* p, const sles_step_params* prm, const void* y, const AdjSrc& mu, void* out,
                int accumulate, int linear, double dcoeff, double* cs_grad, void* work, hipStream_t s);

}  // namespace sles

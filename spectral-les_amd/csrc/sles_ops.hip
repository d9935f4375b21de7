// Pointwise physical/spectral kernels and FP64 reductions behind the
// reference's operator-level call surface (math_formula.py, smagorinsky.py,
// vreman.py, RK4/RK3/ETD2_method.py, SM_solver.py diagnostics).
//
// These are single-pass HBM-bound kernels: one read of every operand and one
// write of every result, grid-stride loops over contiguous elements.  The
// fused time-step path (sles_rhs.hip) does not use them.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <string>

#include "sles.h"
#include "sles_internal.h"

namespace sles {

constexpr int kThreads = 256;

struct Fields {
  void* p[9];
};

static inline unsigned grid_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <typename T>
__device__ __forceinline__ T* fp(const Fields& f, int i) {
  return (T*)f.p[i];
}

// ---------------------------------------------------------------------------
// Physical-space kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_cross(Fields a, Fields b, Fields c, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const T a0 = fp<T>(a, 0)[e], a1 = fp<T>(a, 1)[e], a2 = fp<T>(a, 2)[e];
    const T b0 = fp<T>(b, 0)[e], b1 = fp<T>(b, 1)[e], b2 = fp<T>(b, 2)[e];
    fp<T>(c, 0)[e] = a1 * b2 - a2 * b1;
    fp<T>(c, 1)[e] = a2 * b0 - a0 * b2;
    fp<T>(c, 2)[e] = a0 * b1 - a1 * b0;
  }
}

// a[3*i+j] = d u_j / d x_i  (reference a_{i+1,j+1})
template <typename T>
__global__ void k_smag_nu(Fields a, T* nu, T coeff, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const T a11 = fp<T>(a, 0)[e], a12 = fp<T>(a, 1)[e], a13 = fp<T>(a, 2)[e];
    const T a21 = fp<T>(a, 3)[e], a22 = fp<T>(a, 4)[e], a23 = fp<T>(a, 5)[e];
    const T a31 = fp<T>(a, 6)[e], a32 = fp<T>(a, 7)[e], a33 = fp<T>(a, 8)[e];
    const T s12 = a12 + a21, s13 = a13 + a31, s23 = a23 + a32;
    const T ss = a11 * a11 + T(0.5) * (s12 * s12) + a22 * a22 + T(0.5) * (s13 * s13) +
                 T(0.5) * (s23 * s23) + a33 * a33;
    nu[e] = sqrt(T(2) * ss) * coeff;
  }
}

template <typename T>
__global__ void k_vreman_nu(Fields a, T* nu, T* aa_out, T coeff, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const T a11 = fp<T>(a, 0)[e], a12 = fp<T>(a, 1)[e], a13 = fp<T>(a, 2)[e];
    const T a21 = fp<T>(a, 3)[e], a22 = fp<T>(a, 4)[e], a23 = fp<T>(a, 5)[e];
    const T a31 = fp<T>(a, 6)[e], a32 = fp<T>(a, 7)[e], a33 = fp<T>(a, 8)[e];
    const T b11 = a11 * a11 + a21 * a21 + a31 * a31;
    const T b22 = a12 * a12 + a22 * a22 + a32 * a32;
    const T b33 = a13 * a13 + a23 * a23 + a33 * a33;
    const T b12 = a11 * a12 + a21 * a22 + a31 * a32;
    const T b13 = a11 * a13 + a21 * a23 + a31 * a33;
    const T b23 = a12 * a13 + a22 * a23 + a32 * a33;
    const T sb = sqrt(b11 * b22 - b12 * b12 + b11 * b33 - b13 * b13 + b22 * b33 - b23 * b23);
    const T aa = a11 * a11 + a12 * a12 + a13 * a13 + a21 * a21 + a22 * a22 + a23 * a23 + a31 * a31 +
                 a32 * a32 + a33 * a33;
    // where((aa > 10e-5) * (sb > aa / 10e6), sqrt(1/aa), 0) ; NaN * 0 stays NaN
    const T f = (aa > T(1e-4) && sb > aa / T(1e7)) ? sqrt(T(1) / aa) : T(0);
    nu[e] = sb * (f * coeff);
    if (aa_out) aa_out[e] = aa;
  }
}

template <typename T>
__global__ void k_stress_products(Fields a, const T* nu, Fields tau, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const T v = nu[e];
    fp<T>(tau, 0)[e] = fp<T>(a, 0)[e] * v;
    fp<T>(tau, 1)[e] = T(0.5) * (fp<T>(a, 1)[e] + fp<T>(a, 3)[e]) * v;
    fp<T>(tau, 2)[e] = T(0.5) * (fp<T>(a, 2)[e] + fp<T>(a, 6)[e]) * v;
    fp<T>(tau, 3)[e] = fp<T>(a, 4)[e] * v;
    fp<T>(tau, 4)[e] = T(0.5) * (fp<T>(a, 5)[e] + fp<T>(a, 7)[e]) * v;
    fp<T>(tau, 5)[e] = fp<T>(a, 8)[e] * v;
  }
}

// ---------------------------------------------------------------------------
// Spectral kernels (wavenumbers from indices, never stored)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_curl(Fields u, Fields w, long nc, int n, int nh) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nc; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e, n, nh);
    const T k0 = (T)k.kx, k1 = (T)k.ky, k2 = (T)k.kz;
    const cpx<T> u0 = fp<cpx<T>>(u, 0)[e], u1 = fp<cpx<T>>(u, 1)[e], u2 = fp<cpx<T>>(u, 2)[e];
    // 1j * (K0 a1 - K1 a0) etc. (math_formula.py:22-24)
    fp<cpx<T>>(w, 2)[e] = mul_i(k0 * u1 - k1 * u0);
    fp<cpx<T>>(w, 1)[e] = mul_i(k2 * u0 - k0 * u2);
    fp<cpx<T>>(w, 0)[e] = mul_i(k1 * u2 - k2 * u1);
  }
}

template <typename T>
__global__ void k_gradient(Fields u, Fields g, long nc, int n, int nh) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nc; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e, n, nh);
    const T kk[3] = {(T)k.kx, (T)k.ky, (T)k.kz};
    const cpx<T> uu[3] = {fp<cpx<T>>(u, 0)[e], fp<cpx<T>>(u, 1)[e], fp<cpx<T>>(u, 2)[e]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) fp<cpx<T>>(g, 3 * i + j)[e] = mul_i(kk[i] * uu[j]);  // 1j * K_i * U_j
  }
}

// z * (0 + 1j*s) with numpy's complex product (keeps NaN/inf propagation).
template <typename T>
__device__ __forceinline__ cpx<T> mul_pure_imag(cpx<T> z, T s) {
  return {z.x * T(0) - z.y * s, z.x * s + z.y * T(0)};
}

template <typename T>
__global__ void k_stress_div(Fields th, Fields t, long nc, int n, int nh) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nc; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e, n, nh);
    const T k0 = (T)k.kx, k1 = (T)k.ky, k2 = (T)k.kz;
    const T f = dealias_keep(k, n) ? T(2) : T(0);  // 2.0 * 1j * dealias
    const cpx<T> s11 = fp<cpx<T>>(th, 0)[e], s12 = fp<cpx<T>>(th, 1)[e], s13 = fp<cpx<T>>(th, 2)[e];
    const cpx<T> s22 = fp<cpx<T>>(th, 3)[e], s23 = fp<cpx<T>>(th, 4)[e], s33 = fp<cpx<T>>(th, 5)[e];
    fp<cpx<T>>(t, 0)[e] = mul_pure_imag((k0 * s11 + k1 * s12) + k2 * s13, f);
    fp<cpx<T>>(t, 1)[e] = mul_pure_imag((k0 * s12 + k1 * s22) + k2 * s23, f);
    fp<cpx<T>>(t, 2)[e] = mul_pure_imag((k0 * s13 + k1 * s23) + k2 * s33, f);
  }
}

template <typename T>
__global__ void k_rhs_finish(Fields du, Fields tt, Fields uh, cpx<T>* phat, Fields f1, T nu, T forcing,
                             int has_t, int has_f1, long nc, int n, int nh) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nc; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e, n, nh);
    const int kk[3] = {k.kx, k.ky, k.kz};
    const long k2i = (long)k.kx * k.kx + (long)k.ky * k.ky + (long)k.kz * k.kz;
    const T keep = dealias_keep(k, n) ? T(1) : T(0);
    cpx<T> d[3], t[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      d[i] = keep * fp<cpx<T>>(du, i)[e];  // dU *= dealias
      t[i] = has_t ? fp<cpx<T>>(tt, i)[e] : cpx<T>{T(0), T(0)};
    }
    // P_hat = sum_i (dU + T)_i K_i / K^2   (K_over_K2 = K / max(K^2, 1))
    cpx<T> p = {T(0), T(0)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T kok = (T)kk[i] / (T)(k2i == 0 ? 1 : k2i);
      const cpx<T> s = has_t ? d[i] + t[i] : d[i];
      p = i == 0 ? kok * s : p + kok * s;
    }
    if (phat) phat[e] = p;
    const bool k0 = (k.kx == 0 && k.ky == 0 && k.kz == 0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      d[i] = d[i] - (T)kk[i] * p;  // dU -= P_hat * K
      if (forcing > T(0)) {
        const cpx<T> u = fp<cpx<T>>(uh, i)[e];
        d[i] = d[i] + (k0 ? cpx<T>{T(0), T(0)} : forcing * u);
      }
      if (has_f1) fp<cpx<T>>(f1, i)[e] = d[i];
      if (nu > T(0)) d[i] = d[i] - (nu * (T)k2i) * fp<cpx<T>>(uh, i)[e];
      if (has_t) d[i] = d[i] + t[i];
      fp<cpx<T>>(du, i)[e] = d[i];
    }
  }
}

template <typename T>
__global__ void k_mul_k2(const cpx<T>* in, cpx<T>* out, long nc, int nfields, int n, int nh) {
  const long total = nc * nfields;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e % nc, n, nh);
    const T k2 = (T)((long)k.kx * k.kx + (long)k.ky * k.ky + (long)k.kz * k.kz);
    const cpx<T> v = in[e];
    out[e] = {v.x * k2, v.y * k2};
  }
}

template <typename T>
__global__ void k_scale(cpx<T>* x, long total, T s) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
    x[e] = s * x[e];
}

template <typename T>
__global__ void k_stage(int kind, cpx<T>* uh, const cpx<T>* uh0, cpx<T>* uh1, const cpx<T>* du, T c0, T c1,
                        long total) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const cpx<T> d = du[e];
    if (kind == 0) {
      if (c0 != T(0)) uh[e] = uh0[e] + c0 * d;
      uh1[e] = uh1[e] + c1 * d;
    } else if (kind == 1) {
      uh[e] = uh0[e] + c0 * d;
    } else if (kind == 2) {
      uh[e] = (uh0[e] - c0 * d) + (T(2) * c0) * d;
    } else {
      const cpx<T> s = ((T(1) / T(6)) * d + (T(2) / T(3)) * d) + (T(1) / T(6)) * d;
      uh[e] = uh0[e] + c0 * s;
    }
  }
}

// phi1(z) = (e^z - 1)/z, phi2(z) = (e^z - 1 - z)/z^2; Taylor series for |z| < 1e-2.
__device__ __forceinline__ double phi1d(double z) {
  if (fabs(z) < 1e-2) return 1 + z / 2 * (1 + z / 3 * (1 + z / 4 * (1 + z / 5 * (1 + z / 6))));
  return expm1(z) / z;
}
__device__ __forceinline__ double phi2d(double z) {
  if (fabs(z) < 1e-2) return 0.5 * (1 + z / 3 * (1 + z / 4 * (1 + z / 5 * (1 + z / 6 * (1 + z / 7)))));
  return (expm1(z) - z) / (z * z);
}

template <typename T>
__global__ void k_etd2(int mode, cpx<T>* uh, const cpx<T>* f1, const cpx<T>* f0, double h, double nu, long nc,
                       int n, int nh) {
  const long total = 3 * nc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e % nc, n, nh);
    const long k2i = (long)k.kx * k.kx + (long)k.ky * k.ky + (long)k.kz * k.kz;
    const double c = nu * (double)k2i;
    if (mode == 0) {
      // ETD2_method.py:16-29 literally
      const double ex = exp(-h * c);
      const double a1 = (1 - h * c) + 2 * h * c;
      const double a0 = -ex + 1 - h * c;
      const double den = k2i == 0 ? 1.0 : h * (c * c);
      const cpx<T> u = (T)ex * uh[e];
      const cpx<T> num = (T)a1 * f1[e] + (T)a0 * f0[e];
      uh[e] = u + cpx<T>{num.x / (T)den, num.y / (T)den};
    } else {
      const double z = -h * c;
      const double p1 = phi1d(z), p2 = phi2d(z);
      const T e1 = (T)exp(z), w1 = (T)(h * (p1 + p2)), w0 = (T)(h * p2);
      uh[e] = (e1 * uh[e] + w1 * f1[e]) - w0 * f0[e];
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: fixed grid of kRedBlocks blocks -> partials -> one block.
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double nanmax(double a, double b) { return (isnan(a) || a > b) ? a : b; }

template <int OP>  // 0 sum, 1 max
__device__ double block_reduce(double v) {
  __shared__ double s[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = OP == 0 ? v + w : nanmax(v, w);
  }
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) s[wid] = v;
  block_sync();
  if (wid == 0) {
    v = lane < (int)(blockDim.x / 64) ? s[lane] : (OP == 0 ? 0.0 : -INFINITY);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_xor(v, o, 64);
      v = OP == 0 ? v + w : nanmax(v, w);
    }
  }
  return v;
}

template <typename T, int OP>  // OP 0: sum a^2 (b == nullptr) or sum a*b ; OP 1: max a
__global__ void __launch_bounds__(kThreads) k_reduce_partial(Fields a, Fields b, int nf, int use_b, long n,
                                                             double* partial) {
  double acc = OP == 0 ? 0.0 : -INFINITY;
  for (int f = 0; f < nf; ++f) {
    const T* x = fp<T>(a, f);
    const T* y = use_b ? fp<T>(b, f) : x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
      if (OP == 0)
        acc += (double)x[e] * (double)y[e];
      else
        acc = nanmax(acc, (double)x[e]);
    }
  }
  acc = block_reduce<OP>(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <int OP>
__global__ void __launch_bounds__(kThreads) k_reduce_final(const double* partial, int np, double* out) {
  double acc = OP == 0 ? 0.0 : -INFINITY;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc = OP == 0 ? acc + partial[i] : nanmax(acc, partial[i]);
  acc = block_reduce<OP>(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

constexpr int kMaxBins = 640;  // LDS of the binning kernel: 4 waves x 3 x nbins doubles (60 KiB at most)

// ---------------------------------------------------------------------------
// Binned spectra, deterministic and without atomics.  Modes are visited row
// by row ((kx, ky) rows of N/2+1 kz values); block b owns the rows
// [b rpb, (b+1) rpb) and its wave w every fourth of them, 64 consecutive kz
// per step.  Along a row K^2 grows with kz, so a step's bins are
// non-decreasing across the wave's lanes: a segmented inclusive scan keyed by
// the bin (shuffles) leaves each bin's sum in the last lane of its run, which
// adds it to the wave's own LDS bins.  The block then sums its four waves in
// order into its partial row, and a second kernel sums the partial rows in a
// fixed order: the result does not depend on scheduling.
//   RULE_EK    calculate_Ek (src/SM_solver.py:154-192): bin = isqrt(K^2) / 5,
//              value sum_i |u_i|^2 (numpy.absolute), doubled for kz >= 1
//   RULE_SHELL unit shells |K| in [k - 1/2, k + 1/2), k = 1..nbins (bin k-1),
//              value (1/2) w sum_i |u_i|^2 / N^6, w = 1 on the kz = 0 and
//              kz = N/2 planes and 2 elsewhere (the rfft doubling)
// Partials: [block][3][nbins] doubles (value, sum of |K|, mode count).
// ---------------------------------------------------------------------------
enum { RULE_EK = 0, RULE_SHELL = 1 };
constexpr int kBinWaves = kThreads / 64;
constexpr int kBinMaxBlocks = 1024;

// rows per block of the binned spectra: at least 4 (one per wave), at most
// kBinMaxBlocks blocks
static inline long bin_rows_per_block(int n) {
  const long rows = (long)n * n;
  long r = (rows + kBinMaxBlocks - 1) / kBinMaxBlocks;
  return r < kBinWaves ? kBinWaves : r;
}
static inline long bin_blocks(int n) {
  const long rows = (long)n * n, r = bin_rows_per_block(n);
  return (rows + r - 1) / r;
}

// bin of a mode, or -1 (below the first bin) / INT_MAX (beyond the last):
// both keep the bins of a row non-decreasing in kz
template <int RULE>
__device__ __forceinline__ int mode_bin(long k2, int nbins) {
  if (RULE == RULE_EK) {
    long r = (long)sqrt((double)k2);
    while (r * r > k2) --r;
    while ((r + 1) * (r + 1) <= k2) ++r;
    const long b = r / 5;
    return b < nbins ? (int)b : INT_MAX;
  } else {
    const long sh = (long)floor(sqrt((double)k2) + 0.5);
    if (sh < 1) return -1;
    return sh <= nbins ? (int)(sh - 1) : INT_MAX;
  }
}

template <typename T, int RULE>
__global__ void __launch_bounds__(kThreads) k_bins(const cpx<T>* __restrict__ uh, long nc, int n, int nh, int nbins,
                                                   long rpb, double* __restrict__ part) {
  extern __shared__ double s[];  // [wave][3][nbins]
  for (int i = threadIdx.x; i < kBinWaves * 3 * nbins; i += kThreads) s[i] = 0.0;
  block_sync();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long rows = (long)n * n;
  const long r0 = (long)blockIdx.x * rpb, r1 = r0 + rpb < rows ? r0 + rpb : rows;
  const double c0 = 0.5 / ((double)n * n * n * (double)n * n * n);
  for (long row = r0 + w; row < r1; row += kBinWaves) {
    const int kx = kfreq((int)(row / n), n), ky = kfreq((int)(row % n), n);
    const long q = (long)kx * kx + (long)ky * ky;
    for (int z0 = 0; z0 < nh; z0 += 64) {
      const int kz = z0 + lane;
      const long k2 = q + (long)kz * kz;
      int bin = kz < nh ? mode_bin<RULE>(k2, nbins) : INT_MAX;
      // the rest of the row lies beyond the last bin (bins grow with kz)
      if (__ballot(bin != INT_MAX) == 0) break;
      double v = 0.0, kk = 0.0, c = 0.0;
      if (bin >= 0 && bin < nbins) {
        const long e = row * nh + kz;
        if (RULE == RULE_EK) {
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const cpx<T> u = uh[i * nc + e];
            const double m = hypot(fabs((double)u.x), fabs((double)u.y));  // numpy.absolute
            v += m * m;
          }
          if (kz >= 1) v *= 2.0;
        } else {
          double sq[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const cpx<T> u = uh[i * nc + e];
            sq[i] = (double)u.x * (double)u.x + (double)u.y * (double)u.y;
          }
          const double wgt = (kz == 0 || 2 * kz == n) ? 1.0 : 2.0;
          v = (c0 * wgt) * ((sq[0] + sq[1]) + sq[2]);
        }
        kk = sqrt((double)k2);
        c = 1.0;
      }
      // segmented inclusive scan over the run of equal bins ending at each lane
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const double v2 = __shfl_up(v, d, 64), kk2 = __shfl_up(kk, d, 64), c2 = __shfl_up(c, d, 64);
        const int b2 = __shfl_up(bin, d, 64);
        if (lane >= d && b2 == bin) {
          v += v2;
          kk += kk2;
          c += c2;
        }
      }
      const int bn = __shfl_down(bin, 1, 64);
      if ((lane == 63 || bn != bin) && bin >= 0 && bin < nbins) {
        double* sw = s + (long)w * 3 * nbins;
        sw[bin] += v;
        sw[nbins + bin] += kk;
        sw[2 * nbins + bin] += c;
      }
    }
  }
  block_sync();
  for (int i = threadIdx.x; i < 3 * nbins; i += kThreads) {
    const int qq = i / nbins, b = i % nbins;
    double acc = s[qq * nbins + b];
#pragma unroll
    for (int ww = 1; ww < kBinWaves; ++ww) acc += s[((long)ww * 3 + qq) * nbins + b];
    part[((long)blockIdx.x * 3 + qq) * nbins + b] = acc;
  }
}

// out_q[b] = sum over the partial rows (block q*nbins + b of the grid): each
// thread a fixed stride of rows, then a fixed shuffle tree and the waves in
// order
__global__ void __launch_bounds__(kThreads) k_bins_final(const double* __restrict__ part, int nblocks, int nbins,
                                                         double* o0, double* o1, double* o2) {
  const int qq = blockIdx.x / nbins, b = blockIdx.x % nbins;
  double* out = qq == 0 ? o0 : qq == 1 ? o1 : o2;
  if (!out) return;
  double acc = 0.0;
  for (int r = threadIdx.x; r < nblocks; r += kThreads) acc += part[((long)r * 3 + qq) * nbins + b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  __shared__ double sw[kBinWaves];
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  block_sync();
  if (threadIdx.x == 0) {
    double t = sw[0];
    for (int ww = 1; ww < kBinWaves; ++ww) t += sw[ww];
    out[b] = t;
  }
}

// Vector-Jacobian product of the unit-shell spectrum: g_u = gE[k-1] w u / N^6
// for modes in shells 1..nbins (the gradient of (1/2) w |u|^2 / N^6 with
// respect to (Re u, Im u), as Re + i Im), 0 elsewhere.
template <typename T>
__global__ void k_shell_vjp(const cpx<T>* __restrict__ uh, const double* __restrict__ gE, long nc, int n, int nh,
                            int nbins, cpx<T>* __restrict__ gu) {
  const double c1 = 1.0 / ((double)n * n * n * (double)n * n * n);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nc; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e, n, nh);
    const long k2 = (long)k.kx * k.kx + (long)k.ky * k.ky + (long)k.kz * k.kz;
    const int bin = mode_bin<RULE_SHELL>(k2, nbins);
    double g = 0.0;
    if (bin >= 0 && bin < nbins) g = gE[bin] * ((k.kz == 0 || 2 * k.kz == n) ? 1.0 : 2.0) * c1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const cpx<T> u = uh[i * nc + e];
      gu[i * nc + e] = cpx<T>{(T)(g * (double)u.x), (T)(g * (double)u.y)};
    }
  }
}

static Fields pack(void* const* p, int k) {
  Fields f{};
  for (int i = 0; i < k; ++i) f.p[i] = p ? p[i] : nullptr;
  return f;
}

static bool all_set(void* const* p, int k) {
  if (!p) return false;
  for (int i = 0; i < k; ++i)
    if (!p[i]) return false;
  return true;
}

}  // namespace sles

using namespace sles;

#define SLES_DISPATCH(plan, ...)              \
  do {                                        \
    if ((plan)->precision == SLES_FP64) {     \
      using T = double;                       \
      __VA_ARGS__;                            \
    } else {                                  \
      using T = float;                        \
      __VA_ARGS__;                            \
    }                                         \
  } while (0)

template <int RULE>
static int binned(sles_plan p, const void* uh, int nbins, double* o0, double* o1, double* o2, void* work,
                  hipStream_t s, const char* name) {
  const long nc = ncplx(p);
  const long rpb = bin_rows_per_block(p->n);
  const int nb = (int)bin_blocks(p->n);
  double* part = (double*)work;
  const size_t lds = sizeof(double) * kBinWaves * 3 * (size_t)nbins;
  SLES_DISPATCH(p, hipLaunchKernelGGL((k_bins<T, RULE>), dim3(nb), dim3(kThreads), lds, s, (const cpx<T>*)uh, nc,
                                      p->n, p->nh, nbins, rpb, part));
  hipLaunchKernelGGL(k_bins_final, dim3(3 * nbins), dim3(kThreads), 0, s, (const double*)part, nb, nbins, o0, o1, o2);
  return check_launch(name);
}

extern "C" {

int sles_cross_product(sles_plan p, void* const a[3], void* const b[3], void* const c[3], void* stream) {
  if (!p || !all_set(a, 3) || !all_set(b, 3) || !all_set(c, 3)) return fail("sles_cross_product: null argument");
  const long n = nreal(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_cross<T>, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(a, 3), pack(b, 3), pack(c, 3), n));
  return check_launch("cross_product");
}

int sles_spectral_curl(sles_plan p, void* const u[3], void* const w[3], void* stream) {
  if (!p || !all_set(u, 3) || !all_set(w, 3)) return fail("sles_spectral_curl: null argument");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_curl<T>, dim3(grid_for(nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(u, 3), pack(w, 3), nc, p->n, p->nh));
  return check_launch("spectral_curl");
}

int sles_spectral_gradient(sles_plan p, void* const u[3], void* const g[9], void* stream) {
  if (!p || !all_set(u, 3) || !all_set(g, 9)) return fail("sles_spectral_gradient: null argument");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_gradient<T>, dim3(grid_for(nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(u, 3), pack(g, 9), nc, p->n, p->nh));
  return check_launch("spectral_gradient");
}

int sles_smagorinsky_nu(sles_plan p, void* const a[9], void* nu, double coeff, void* stream) {
  if (!p || !all_set(a, 9) || !nu) return fail("sles_smagorinsky_nu: null argument");
  const long n = nreal(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_smag_nu<T>, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(a, 9), (T*)nu, (T)coeff, n));
  return check_launch("smagorinsky_nu");
}

int sles_vreman_nu(sles_plan p, void* const a[9], void* nu, void* aa, double coeff, void* stream) {
  if (!p || !all_set(a, 9) || !nu) return fail("sles_vreman_nu: null argument");
  const long n = nreal(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_vreman_nu<T>, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(a, 9), (T*)nu, (T*)aa, (T)coeff, n));
  return check_launch("vreman_nu");
}

int sles_stress_products(sles_plan p, void* const a[9], const void* nu, void* const tau[6], void* stream) {
  if (!p || !all_set(a, 9) || !nu || !all_set(tau, 6)) return fail("sles_stress_products: null argument");
  const long n = nreal(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_stress_products<T>, dim3(grid_for(n)), dim3(kThreads), 0,
                                      (hipStream_t)stream, pack(a, 9), (const T*)nu, pack(tau, 6), n));
  return check_launch("stress_products");
}

int sles_stress_divergence(sles_plan p, void* const th[6], void* const t[3], void* stream) {
  if (!p || !all_set(th, 6) || !all_set(t, 3)) return fail("sles_stress_divergence: null argument");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_stress_div<T>, dim3(grid_for(nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(th, 6), pack(t, 3), nc, p->n, p->nh));
  return check_launch("stress_divergence");
}

int sles_rhs_finish(sles_plan p, void* const du[3], void* const t[3], void* const uh[3], void* phat,
                    void* const f1[3], double nu, double forcing, void* stream) {
  if (!p || !all_set(du, 3) || !all_set(uh, 3)) return fail("sles_rhs_finish: null argument");
  const int has_t = t != nullptr && all_set(t, 3);
  const int has_f1 = f1 != nullptr && all_set(f1, 3);
  if ((t && !has_t) || (f1 && !has_f1)) return fail("sles_rhs_finish: partially null field array");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_rhs_finish<T>, dim3(grid_for(nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      pack(du, 3), pack(t, has_t ? 3 : 0), pack(uh, 3), (cpx<T>*)phat,
                                      pack(f1, has_f1 ? 3 : 0), (T)nu, (T)forcing, has_t, has_f1, nc, p->n,
                                      p->nh));
  return check_launch("rhs_finish");
}

int sles_mul_k2(sles_plan p, const void* in, void* out, int nfields, void* stream) {
  if (!p || !in || !out || nfields < 1) return fail("sles_mul_k2: bad argument");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_mul_k2<T>, dim3(grid_for(nc * nfields)), dim3(kThreads), 0,
                                      (hipStream_t)stream, (const cpx<T>*)in, (cpx<T>*)out, nc, nfields, p->n,
                                      p->nh));
  return check_launch("mul_k2");
}

int sles_scale(sles_plan p, void* x, int nfields, double s, void* stream) {
  if (!p || !x || nfields < 1) return fail("sles_scale: bad argument");
  const long total = (long)nfields * ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(total)), dim3(kThreads), 0, (hipStream_t)stream,
                                      (cpx<T>*)x, total, (T)s));
  return check_launch("scale");
}

int sles_stage_update(sles_plan p, int kind, void* uh, const void* uh0, void* uh1, const void* du, double c0,
                      double c1, void* stream) {
  if (!p || !uh || !du) return fail("sles_stage_update: null argument");
  if (kind < 0 || kind > 3) return fail("sles_stage_update: kind must be 0..3");
  if (kind == 0 && !uh1) return fail("sles_stage_update: kind 0 needs u_hat1");
  if ((kind != 0 || c0 != 0.0) && !uh0) return fail("sles_stage_update: u_hat0 required");
  const long total = 3 * ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_stage<T>, dim3(grid_for(total)), dim3(kThreads), 0, (hipStream_t)stream,
                                      kind, (cpx<T>*)uh, (const cpx<T>*)uh0, (cpx<T>*)uh1, (const cpx<T>*)du,
                                      (T)c0, (T)c1, total));
  return check_launch("stage_update");
}

int sles_etd2_update(sles_plan p, int mode, void* uh, const void* f1, const void* f0, double dt, double nu,
                     void* stream) {
  if (!p || !uh || !f1 || !f0) return fail("sles_etd2_update: null argument");
  if (mode != 0 && mode != 1) return fail("sles_etd2_update: mode must be 0 or 1");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_etd2<T>, dim3(grid_for(3 * nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      mode, (cpx<T>*)uh, (const cpx<T>*)f1, (const cpx<T>*)f0, dt, nu, nc, p->n,
                                      p->nh));
  return check_launch("etd2_update");
}

size_t sles_reduce_work_bytes(void) { return sizeof(double) * kRedBlocks; }

static int reduce_common(sles_plan p, int op, void* const* a, void* const* b, int nf, double* out, void* work,
                         void* stream, const char* name) {
  if (!p || !out || !work || nf < 1 || nf > 9 || !all_set(a, nf) || (b && !all_set(b, nf)))
    return fail(std::string(name) + ": bad argument");
  const long n = nreal(p);
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)work;
  if (op == 0) {
    SLES_DISPATCH(p, hipLaunchKernelGGL((k_reduce_partial<T, 0>), dim3(kRedBlocks), dim3(kThreads), 0, s,
                                        pack(a, nf), pack(b, b ? nf : 0), nf, b ? 1 : 0, n, part));
    hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(kThreads), 0, s, (const double*)part, kRedBlocks, out);
  } else {
    SLES_DISPATCH(p, hipLaunchKernelGGL((k_reduce_partial<T, 1>), dim3(kRedBlocks), dim3(kThreads), 0, s,
                                        pack(a, nf), pack(b, 0), nf, 0, n, part));
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kThreads), 0, s, (const double*)part, kRedBlocks, out);
  }
  return check_launch(name);
}

int sles_sum_squares(sles_plan p, void* const f[], int nf, double* out, void* work, void* stream) {
  return reduce_common(p, 0, f, nullptr, nf, out, work, stream, "sles_sum_squares");
}

int sles_dot(sles_plan p, void* const a[], void* const b[], int nf, double* out, void* work, void* stream) {
  if (!b) return fail("sles_dot: null argument");
  return reduce_common(p, 0, a, b, nf, out, work, stream, "sles_dot");
}

int sles_max(sles_plan p, void* const f[], int nf, double* out, void* work, void* stream) {
  return reduce_common(p, 1, f, nullptr, nf, out, work, stream, "sles_max");
}

size_t sles_spectrum_work_bytes(sles_plan p, int nbins) {
  if (!p || nbins < 1) return 0;
  return sizeof(double) * 3 * (size_t)nbins * (size_t)bin_blocks(p->n);
}

int sles_spectrum(sles_plan p, const void* uh, int nbins, double* ek, double* ks, double* cnt, void* work,
                  void* stream) {
  if (!p || !uh || !ek || !ks || !cnt || !work) return fail("sles_spectrum: null argument");
  if (nbins < 1 || nbins > kMaxBins) return fail("sles_spectrum: nbins must be in [1, 640]");
  return binned<RULE_EK>(p, uh, nbins, ek, ks, cnt, work, (hipStream_t)stream, "spectrum");
}

int sles_shell_spectrum(sles_plan p, const void* uh, int kmax, double* e, void* work, void* stream) {
  if (!p || !uh || !e || !work) return fail("sles_shell_spectrum: null argument");
  if (kmax < 1 || kmax > kMaxBins) return fail("sles_shell_spectrum: kmax must be in [1, 640]");
  return binned<RULE_SHELL>(p, uh, kmax, e, nullptr, nullptr, work, (hipStream_t)stream, "shell_spectrum");
}

int sles_shell_spectrum_vjp(sles_plan p, const void* uh, int kmax, const double* grad_e, void* grad_u,
                            void* stream) {
  if (!p || !uh || !grad_e || !grad_u) return fail("sles_shell_spectrum_vjp: null argument");
  if (kmax < 1 || kmax > kMaxBins) return fail("sles_shell_spectrum_vjp: kmax must be in [1, 640]");
  const long nc = ncplx(p);
  SLES_DISPATCH(p, hipLaunchKernelGGL(k_shell_vjp<T>, dim3(grid_for(nc)), dim3(kThreads), 0, (hipStream_t)stream,
                                      (const cpx<T>*)uh, grad_e, nc, p->n, p->nh, kmax, (cpx<T>*)grad_u));
  return check_launch("shell_spectrum_vjp");
}

}  // extern "C"

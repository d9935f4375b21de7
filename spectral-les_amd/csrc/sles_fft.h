// Stockham FFT building blocks for gfx950 (CDNA4, wave64).
//
// A "line" is one 1D transform of length N held by NT = N/P threads; thread t
// of a line holds the P points t + m*NT (m = 0..P-1) in registers.  That
// distribution is the same on input and on output of every transform, so
// pointwise work between an inverse and a forward transform needs no data
// exchange, and global loads/stores along a contiguous axis are coalesced
// (consecutive t -> consecutive addresses).
//
// Stage s of radix R (Ns = product of the earlier radices) runs N/R
// butterflies; butterfly j = t + q*NT (q = 0..P/R-1) of thread t reads the
// slots q + r*P/R, multiplies by W^(r*k*N/(Ns*R)) with k = j mod Ns, runs a
// radix-R DFT and sends output r to position (j-k)*R + k + r*Ns.  Between
// stages the outputs go through LDS; after the last stage every output is
// already in the slot it started from.  (Stockham autosort, e.g. Govindaraju
// et al., "High performance discrete Fourier transforms on graphics
// processors", SC'08.)
#pragma once

#include <hip/hip_runtime.h>

namespace sles {

// Block barrier that is also a compiler barrier for memory operations.
// __syncthreads() alone (fence + s_barrier + fence) has let the scheduler sink
// LDS reads of an exchange below the barrier that protects the next writes
// (observed on gfx950 with FP32 lines), which races between waves.  The
// "memory" clobbers keep every load and store on its side of the barrier.
__device__ __forceinline__ void block_sync() {
  asm volatile("" ::: "memory");
  __syncthreads();
  asm volatile("" ::: "memory");
}

// Synchronisation of one wavefront's LDS traffic, for exchanges whose lines
// all live inside one wavefront.  LDS operations of a wavefront execute in
// program order, so only the compiler has to be kept from moving loads and
// stores across this point (wavefront-scope fences plus "memory" clobbers);
// no s_barrier is issued and the other waves of the block run on.
__device__ __forceinline__ void wave_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}
template <bool WAVE>
__device__ __forceinline__ void line_sync() {
  if constexpr (WAVE)
    wave_sync();
  else
    block_sync();
}

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

// Element idx of an array whose base pointer is the same for the whole wave.
// The byte offset is formed in 32 bits (every field addressed this way is
// below 4 GiB), so the access can use the scalar base register plus one
// 32-bit vector offset instead of a 64-bit address per element.
template <typename T>
__device__ __forceinline__ T ld32(const T* base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st32(T* base, unsigned idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)) = v;
}

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }
// i * a
template <typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return {-a.y, a.x}; }
// -i * a
template <typename T>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) { return {a.y, -a.x}; }
// (i * s) * a for real s: the spectral derivative factor i*K.
template <typename T>
__device__ __forceinline__ cpx<T> mul_is(T s, cpx<T> a) { return {-s * a.y, s * a.x}; }

// ---------------------------------------------------------------------------
// Line plans: radices per stage and points per thread P (every radix divides P).
// ---------------------------------------------------------------------------
template <int N_, int P_, int... Rs>
struct Plan {
  static constexpr int N = N_;
  static constexpr int P = P_;
  static constexpr int NT = N_ / P_;
  template <template <int...> class F>
  using apply = F<Rs...>;
};

// Default plan per length: radix-8 where possible (fewest LDS exchanges).
template <int N>
struct LinePlan;
#define SLES_PLAN(NN, PP, ...) \
  template <>                  \
  struct LinePlan<NN> : Plan<NN, PP, __VA_ARGS__> {};
SLES_PLAN(8, 8, 8)
SLES_PLAN(12, 12, 4, 3)
SLES_PLAN(16, 4, 4, 4)
SLES_PLAN(24, 12, 2, 3, 4)
SLES_PLAN(32, 8, 8, 4)
SLES_PLAN(48, 12, 4, 4, 3)
SLES_PLAN(64, 8, 8, 8)
SLES_PLAN(96, 12, 3, 4, 4, 2)
SLES_PLAN(128, 8, 8, 4, 4)
SLES_PLAN(192, 12, 4, 4, 4, 3)
SLES_PLAN(256, 8, 8, 8, 4)
SLES_PLAN(384, 12, 4, 4, 4, 3, 2)
SLES_PLAN(512, 8, 8, 8, 8)
#undef SLES_PLAN

// Alternative plans with fewer points per thread (selected per pass by the
// tuning knobs): radix-4 (4 points per thread: half the registers of the
// radix-8 plans, twice the threads per line, more LDS exchanges) for powers
// of two, and radix 3 then radix 2 (6 points per thread, half the registers
// of the 12-point plans, which spill in the fused passes) for the sizes with
// a factor 3.
template <int N>
struct LinePlan4 : LinePlan<N> {};
#define SLES_PLAN4(NN, ...) \
  template <>               \
  struct LinePlan4<NN> : Plan<NN, 4, __VA_ARGS__> {};
SLES_PLAN4(32, 4, 4, 2)
SLES_PLAN4(64, 4, 4, 4)
SLES_PLAN4(128, 4, 4, 4, 2)
SLES_PLAN4(256, 4, 4, 4, 4)
SLES_PLAN4(512, 4, 4, 4, 4, 2)
#undef SLES_PLAN4
#define SLES_PLAN6(NN, ...) \
  template <>               \
  struct LinePlan4<NN> : Plan<NN, 6, __VA_ARGS__> {};
SLES_PLAN6(12, 3, 2, 2)
SLES_PLAN6(24, 3, 2, 2, 2)
SLES_PLAN6(48, 3, 2, 2, 2, 2)
SLES_PLAN6(96, 3, 2, 2, 2, 2, 2)
SLES_PLAN6(192, 3, 2, 2, 2, 2, 2, 2)
SLES_PLAN6(384, 3, 2, 2, 2, 2, 2, 2, 2)
#undef SLES_PLAN6

// ---------------------------------------------------------------------------
// Radix-R DFT in registers.  DIR = -1 forward (e^{-2 pi i nk/R}), +1 inverse.
// ---------------------------------------------------------------------------
template <typename T, int DIR>
__device__ __forceinline__ cpx<T> rot_q(cpx<T> a) {  // multiply by e^{DIR i pi/2}
  return DIR < 0 ? mul_mi(a) : mul_i(a);
}

template <typename T, int R, int DIR>
struct Butterfly;

template <typename T, int DIR>
struct Butterfly<T, 2, DIR> {
  __device__ __forceinline__ static void run(cpx<T>* a) {
    cpx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <typename T, int DIR>
struct Butterfly<T, 3, DIR> {
  __device__ __forceinline__ static void run(cpx<T>* a) {
    const T s = T(0.86602540378443864676372317075293618);  // sin(2 pi/3)
    cpx<T> sum = a[1] + a[2];
    cpx<T> dif = a[1] - a[2];
    cpx<T> m = a[0] - T(0.5) * sum;
    // forward: -i s (a1 - a2); inverse: +i s (a1 - a2)
    cpx<T> u = DIR < 0 ? cpx<T>{s * dif.y, -s * dif.x} : cpx<T>{-s * dif.y, s * dif.x};
    a[0] = a[0] + sum;
    a[1] = m + u;
    a[2] = m - u;
  }
};

template <typename T, int DIR>
struct Butterfly<T, 4, DIR> {
  __device__ __forceinline__ static void run(cpx<T>* a) {
    cpx<T> t0 = a[0] + a[2];
    cpx<T> t1 = a[0] - a[2];
    cpx<T> t2 = a[1] + a[3];
    cpx<T> t3 = rot_q<T, DIR>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  }
};

template <typename T, int DIR>
struct Butterfly<T, 8, DIR> {
  __device__ __forceinline__ static void run(cpx<T>* a) {
    const T h = T(0.70710678118654752440084436210484904);  // 1/sqrt(2)
    cpx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cpx<T> o[4] = {a[1], a[3], a[5], a[7]};
    Butterfly<T, 4, DIR>::run(e);
    Butterfly<T, 4, DIR>::run(o);
    // W8^k = e^{DIR i pi k/4}
    cpx<T> o1 = DIR < 0 ? cpx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)}
                        : cpx<T>{h * (o[1].x - o[1].y), h * (o[1].y + o[1].x)};
    cpx<T> o2 = rot_q<T, DIR>(o[2]);
    cpx<T> o3 = DIR < 0 ? cpx<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)}
                        : cpx<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)};
    a[0] = e[0] + o[0];
    a[4] = e[0] - o[0];
    a[1] = e[1] + o1;
    a[5] = e[1] - o1;
    a[2] = e[2] + o2;
    a[6] = e[2] - o2;
    a[3] = e[3] + o3;
    a[7] = e[3] - o3;
  }
};

// ---------------------------------------------------------------------------
// LDS index maps used for the exchange between stages.
// ---------------------------------------------------------------------------
// Lines stored contiguously (contiguous-axis passes): position p of the line.
// LDS serves 128 bytes per clock from 32 four-byte banks, so E = 8 (FP64) /
// 16 (FP32) complex elements share one bank row.  When E divides N, the
// elements of each row are permuted by an XOR with g(row): p ^ g(p / E),
//   g(b) = (b ^ 2b) mod 8                    (FP64)
//   g(b) = (b ^ 2b ^ 4b) mod 16              (FP32)
// which keeps every access group of an exchange (stride-R writes, the
// R-blocks of the second stage, contiguous reads) on distinct banks for the
// radix-4 and radix-8 plans of 64..512 points.  Other N keep the older
// layout: one pad element after every 256 bytes.  A line occupies STRIDE
// elements.  plain() is the identity layout of the same region, for the
// two-for-one split, whose mirrored reads (N - k) are conflict-free only
// without a permutation.
struct LdsPlain {
  int base;
  __device__ __forceinline__ int operator()(int p) const { return base + p; }
};
template <typename T, int N>
struct LdsLine {
  static constexpr int LOGE = sizeof(T) == 8 ? 3 : 4;
  static constexpr int E = 1 << LOGE;
  static constexpr bool SWZ = N % E == 0;
  static constexpr int STRIDE = SWZ ? N : N + (N >> (LOGE + 1));
  int base;
  __device__ __forceinline__ int operator()(int p) const {
    if constexpr (SWZ) {
      const int b = p >> LOGE;
      const int g = (sizeof(T) == 8 ? (b ^ (b << 1)) : (b ^ (b << 1) ^ (b << 2))) & (E - 1);
      return base + (p ^ g);
    } else {
      return base + p + (p >> (LOGE + 1));
    }
  }
  __device__ __forceinline__ LdsPlain plain() const { return {base}; }
};
// Columns interleaved (strided passes): position p of column c, TK columns.
template <int TK>
struct LdsColumn {
  int c;
  __device__ __forceinline__ int operator()(int p) const { return p * TK + c; }
};

// ---------------------------------------------------------------------------
// Twiddle sources.  tw[e] = exp(-2 pi i e / N) (forward table; the inverse
// conjugates it).  Stage (Ns, R) of thread t needs, for each of its Q = P/R
// butterflies j = t + q*NT and r = 1..R-1, the factor tw[r * (j mod Ns) * N/(Ns R)].
//   TwTable: read from the table (L1/L2-resident) at every stage.
//   TwRegs:  every factor of every stage read once into registers (kernels
//            that run many transforms per thread reuse them).
// OFF is the running index of the stage's first factor in TwRegs.
// ---------------------------------------------------------------------------
template <typename T>
struct TwTable {
  const cpx<T>* __restrict__ tw;
  template <int N, int NT, int Ns, int R, int OFF>
  __device__ __forceinline__ cpx<T> get(int q, int r, int t) const {
    const int k = (t + q * NT) % Ns;
    return tw[r * k * (N / (Ns * R))];
  }
};

// Table twiddles addressed as a uniform base plus a 32-bit offset (one VGPR
// per factor address instead of two), for kernels at the register limit.
template <typename T>
struct TwTable32 {
  const cpx<T>* __restrict__ tw;
  template <int N, int NT, int Ns, int R, int OFF>
  __device__ __forceinline__ cpx<T> get(int q, int r, int t) const {
    const int k = (t + q * NT) % Ns;
    return ld32(tw, (unsigned)(r * k * (N / (Ns * R))));
  }
};

// Number of register twiddles of a plan (stages after the first).
template <int P, int Ns, int... Rs>
struct TwCount;
template <int P, int Ns>
struct TwCount<P, Ns> {
  static constexpr int value = 0;
};
template <int P, int Ns, int R, int... Rest>
struct TwCount<P, Ns, R, Rest...> {
  static constexpr int value = (Ns > 1 ? (P / R) * (R - 1) : 0) + TwCount<P, Ns * R, Rest...>::value;
};

template <typename T, int N, int P, int Ns, int OFF, int... Rs>
struct TwLoad;
template <typename T, int N, int P, int Ns, int OFF>
struct TwLoad<T, N, P, Ns, OFF> {
  __device__ __forceinline__ static void run(cpx<T>*, const cpx<T>* __restrict__, int) {}
};
template <typename T, int N, int P, int Ns, int OFF, int R, int... Rest>
struct TwLoad<T, N, P, Ns, OFF, R, Rest...> {
  static constexpr int CNT = Ns > 1 ? (P / R) * (R - 1) : 0;
  __device__ __forceinline__ static void run(cpx<T>* w, const cpx<T>* __restrict__ tw, int t) {
    if constexpr (Ns > 1) {
      constexpr int NT = N / P, Q = P / R;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k = (t + q * NT) % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) w[OFF + q * (R - 1) + (r - 1)] = tw[r * k * (N / (Ns * R))];
      }
    }
    TwLoad<T, N, P, Ns * R, OFF + CNT, Rest...>::run(w, tw, t);
  }
};

template <typename T, class PL>
struct TwRegs {
  template <int... Rs>
  using count_of = TwCount<PL::P, 1, Rs...>;
  static constexpr int COUNT = PL::template apply<count_of>::value;
  template <int... Rs>
  using loader = TwLoad<T, PL::N, PL::P, 1, 0, Rs...>;
  cpx<T> w[COUNT > 0 ? COUNT : 1];
  __device__ __forceinline__ TwRegs(const cpx<T>* __restrict__ tw, int t) {
    PL::template apply<loader>::run(w, tw, t);
  }
  template <int N, int NT, int Ns, int R, int OFF>
  __device__ __forceinline__ cpx<T> get(int q, int r, int) const {
    return w[OFF + q * (R - 1) + (r - 1)];
  }
};

// Register twiddles when REG, else the table pointer (for kernels that are
// already near the register limit).
template <typename T, class PL, bool REG>
struct TwPick {
  using type = TwRegs<T, PL>;
  __device__ __forceinline__ static type make(const cpx<T>* __restrict__ tw, int t) { return type(tw, t); }
};
template <typename T, class PL>
struct TwPick<T, PL, false> {
  using type = const cpx<T>*;
  __device__ __forceinline__ static type make(const cpx<T>* __restrict__ tw, int) { return tw; }
};

// ---------------------------------------------------------------------------
// One stage: twiddle + butterflies on the P register slots of thread t.
// ---------------------------------------------------------------------------
template <typename T, int N, int P, int DIR, int R, int Ns, int OFF, class TW>
__device__ __forceinline__ void line_stage(cpx<T> (&v)[P], int t, const TW& tw) {
  constexpr int NT = N / P;
  constexpr int Q = P / R;
  static_assert(P % R == 0, "radix must divide points per thread");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cpx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + r * Q];
    if constexpr (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = tw.template get<N, NT, Ns, R, OFF>(q, r, t);
        if constexpr (DIR > 0) w.y = -w.y;
        a[r] = cmul(a[r], w);
      }
    }
    Butterfly<T, R, DIR>::run(a);
#pragma unroll
    for (int r = 0; r < R; ++r) v[q + r * Q] = a[r];
  }
}

// TRAIL: end with a second barrier, so the buffer can be written again at
// once.  Without it the next write must go to another buffer (ping-pong).
template <typename T, int N, int P, int R, int Ns, bool WAVE, bool TRAIL, class Map>
__device__ __forceinline__ void line_exchange(cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t) {
  constexpr int NT = N / P;
  constexpr int Q = P / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = t + q * NT;
    const int k = j % Ns;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[map(base + r * Ns)] = v[q + r * Q];
  }
  line_sync<WAVE>();
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = lds[map(t + m * NT)];
  if constexpr (TRAIL) line_sync<WAVE>();
}

// PPB > 0: exchanges alternate between lds and lds + PPB (two buffers of PPB
// elements), so only the last exchange needs its trailing barrier: a thread
// writes a buffer again only after the barrier of the exchange in between,
// which every thread reaches after its reads of that buffer.  EX: index of
// the next exchange.
template <typename T, int N, int P, int DIR, bool WAVE, int PPB, int Ns, int OFF, int EX, int... Rs>
struct LineStages;

template <typename T, int N, int P, int DIR, bool WAVE, int PPB, int Ns, int OFF, int EX, int R>
struct LineStages<T, N, P, DIR, WAVE, PPB, Ns, OFF, EX, R> {
  template <class Map, class TW>
  __device__ __forceinline__ static void run(cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t, const TW& tw) {
    line_stage<T, N, P, DIR, R, Ns, OFF>(v, t, tw);
  }
};

template <typename T, int N, int P, int DIR, bool WAVE, int PPB, int Ns, int OFF, int EX, int R, int R2, int... Rest>
struct LineStages<T, N, P, DIR, WAVE, PPB, Ns, OFF, EX, R, R2, Rest...> {
  static constexpr int NEXT = OFF + (Ns > 1 ? (P / R) * (R - 1) : 0);
  static constexpr bool LAST = sizeof...(Rest) == 0;
  template <class Map, class TW>
  __device__ __forceinline__ static void run(cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t, const TW& tw) {
    line_stage<T, N, P, DIR, R, Ns, OFF>(v, t, tw);
    cpx<T>* const buf = (PPB > 0 && (EX & 1)) ? lds + PPB : lds;
    line_exchange<T, N, P, R, Ns, WAVE, PPB == 0 || LAST>(v, buf, map, t);
    LineStages<T, N, P, DIR, WAVE, PPB, Ns * R, NEXT, EX + 1, R2, Rest...>::run(v, lds, map, t, tw);
  }
};

// WAVE: every line of the block lies inside one wavefront (NT <= 64 and
// 64 % NT == 0, lines packed by thread index), so the exchanges need only
// wavefront synchronisation.
// PPB > 0: lds holds two buffers of PPB elements each (see LineStages).
template <typename T, class PL, int DIR, bool WAVE = false, int PPB = 0>
struct LineFFTP {
  static constexpr int N = PL::N;
  static constexpr int P = PL::P;
  static constexpr int NT = PL::NT;
  static_assert(!WAVE || (NT <= 64 && 64 % NT == 0), "wavefront-local lines need NT | 64");
  template <int... Rs>
  using stages = LineStages<T, N, P, DIR, WAVE, PPB, 1, 0, 0, Rs...>;
  using S = typename PL::template apply<stages>;
  // Transform the line held by thread t (0 <= t < NT).  Every thread of the
  // block (WAVE: of the wavefront) must call this: it contains barriers when
  // the plan has more than one stage.  tw: the twiddle table, or a
  // TwRegs<T, PL> of thread t.
  template <class Map>
  __device__ __forceinline__ static void run(cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t,
                                             const cpx<T>* __restrict__ tw) {
    S::run(v, lds, map, t, TwTable<T>{tw});
  }
  template <class Map>
  __device__ __forceinline__ static void run(cpx<T> (&v)[P], cpx<T>* lds, const Map& map, int t,
                                             const TwRegs<T, PL>& tw) {
    S::run(v, lds, map, t, tw);
  }
};

// Lines of this plan fit inside one wavefront.
template <class PL>
struct WaveLocal {
  static constexpr bool value = PL::NT <= 64 && 64 % PL::NT == 0;
};

template <typename T, int N, int DIR>
struct LineFFT : LineFFTP<T, LinePlan<N>, DIR> {};

}  // namespace sles

// Measurement kernel: the memory access pattern of the strided passes alone.
// A tile of TK consecutive columns of each (N rows) x (M = N (N/2+1)
// columns) spectral field, rows `plane` elements apart (the plan's workspace
// layout: M + the SLES_TUNE_PAD padding); thread (c = tid % TK, t = tid / TK)
// reads rows t + m NT (m < N/NT) of every input field, sums them, and writes
// the sum to every output field: the loads and stores of an x pass with no
// transform.  Its bandwidth is the ceiling of a pass with that tile shape.
#include <hip/hip_runtime.h>

#include "sles.h"
#include "sles_internal.h"

namespace sles {

template <typename T, int N, int TK, int NT>
__global__ void __launch_bounds__(TK* NT)
    k_stream(const cpx<T>* __restrict__ in, int nin, cpx<T>* __restrict__ out, int nout, long plane, int xcd) {
  constexpr int P = N / NT;
  const long C = N * plane;
  const int c = threadIdx.x % TK, t = threadIdx.x / TK;
  const long g = tile_id(xcd) * TK + c;
  cpx<T> acc[P];
#pragma unroll
  for (int m = 0; m < P; ++m) acc[m] = {T(0), T(0)};
#pragma unroll 1
  for (int f = 0; f < nin; ++f)
#pragma unroll
    for (int m = 0; m < P; ++m) acc[m] = acc[m] + in[f * C + (long)(t + m * NT) * plane + g];
#pragma unroll 1
  for (int f = 0; f < nout; ++f)
#pragma unroll
    for (int m = 0; m < P; ++m) out[f * C + (long)(t + m * NT) * plane + g] = acc[m];
}

template <typename T, int N, int TK, int NT>
static int launch_stream(const sles_plan_s* p, const void* in, int nin, void* out, int nout, hipStream_t s) {
  static_assert(N % NT == 0 && ((long)N * (N / 2 + 1)) % TK == 0, "tile must divide the field");
  const unsigned grid = (unsigned)((long)N * (N / 2 + 1) / TK);
  hipLaunchKernelGGL((k_stream<T, N, TK, NT>), dim3(grid), dim3(TK * NT), 0, s, (const cpx<T>*)in, nin,
                     (cpx<T>*)out, nout, plane_w(p), p->tune_xcd);
  return check_launch("stream probe");
}

// tile (columns x threads per column): 0 the radix-4 x-forward tile (FP64
// 8 x 64, FP32 16 x 64); 1 16-column tiles (16 x 32, FP32 at 512 16 x 64:
// the radix-8 strided shape); 2 32 x 16
template <typename T, int N>
static int stream_t(const sles_plan_s* p, const void* in, int nin, void* out, int nout, int tile, hipStream_t s) {
  constexpr bool D = sizeof(T) == 8;
  switch (tile) {
    case 0: return launch_stream<T, N, D ? 8 : 16, 64>(p, in, nin, out, nout, s);
    case 1: return launch_stream<T, N, 16, N == 512 && !D ? 64 : 32>(p, in, nin, out, nout, s);
    case 2: return launch_stream<T, N, 32, 16>(p, in, nin, out, nout, s);
    default: return fail("sles_stream_probe: tile must be 0, 1 or 2");
  }
}

}  // namespace sles

using namespace sles;

extern "C" int sles_stream_probe(sles_plan p, const void* in, int nin, void* out, int nout, int tile, void* stream) {
  if (!p || (nin > 0 && !in) || (nout > 0 && !out)) return fail("sles_stream_probe: null argument");
  if (nin < 0 || nout < 0) return fail("sles_stream_probe: field counts must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  if (p->n == 256)
    return p->precision == SLES_FP64 ? stream_t<double, 256>(p, in, nin, out, nout, tile, s)
                                     : stream_t<float, 256>(p, in, nin, out, nout, tile, s);
  if (p->n == 512)
    return p->precision == SLES_FP64 ? stream_t<double, 512>(p, in, nin, out, nout, tile, s)
                                     : stream_t<float, 512>(p, in, nin, out, nout, tile, s);
  return fail("sles_stream_probe: N = 256 and 512 only");
}

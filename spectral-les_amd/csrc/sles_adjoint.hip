// C ABI of the hand-written adjoint RHS (kernels in sles_adjoint.h) and the
// small vector helpers the adjoint time stepping uses.
#include "sles_adjoint.h"

namespace sles {

// out = a x + b y (out may be y)
template <typename T>
__global__ void k_axpby(long n, double a, const cpx<T>* x, double b, const cpx<T>* y, cpx<T>* out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = (T)a * x[e] + (T)b * y[e];
}

// out = t_0 + t_1 + ... + t_{n-1}, summed in that order (out may be a term)
struct SumTerms {
  const void* t[8];
  int n;
};
template <typename T>
__global__ void k_sum_fields(long n, SumTerms st, cpx<T>* out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    cpx<T> r = ((const cpx<T>*)st.t[0])[e];
    for (int i = 1; i < st.n; ++i) r = r + ((const cpx<T>*)st.t[i])[e];
    out[e] = r;
  }
}

template <typename T>
__global__ void k_etd2_adj(long nc, int n, int nh, double h, double nu, const cpx<T>* l1, const cpx<T>* l2,
                           cpx<T>* mu, cpx<T>* lo) {
  const long total = 3 * nc;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const KIndex k = kindex(e % nc, n, nh);
    const long k2 = (long)k.kx * k.kx + (long)k.ky * k.ky + (long)k.kz * k.kz;
    const double z = -h * nu * (double)k2;
    const double p1 = phi1_d(z), p2 = phi2_d(z);
    const cpx<T> a = l1[e];
    const cpx<T> b = l2 ? l2[e] : cpx<T>{T(0), T(0)};
    mu[e] = (T)(h * (p1 + p2)) * a - (T)(h * p2) * b;
    lo[e] = (T)exp(z) * a;
  }
}

}  // namespace sles

using namespace sles;

extern "C" {

size_t sles_adjoint_work_bytes(sles_plan p, int les_model) {
  (void)les_model;
  if (!p) return 0;
  const size_t esz = p->precision == SLES_FP64 ? 16 : 8;
  return kAdjFields * (size_t)ncplx(p) * esz + sizeof(double) * ((size_t)p->n * p->n + 64);
}

int sles_adjoint_rhs(sles_plan p, const sles_step_params* prm, const void* y, const void* mu, void* out,
                     int accumulate, int linear, double dcoeff, double* cs_grad, void* work, void* stream) {
  if (!p || !prm || !y || !mu || !out || !work) return fail("sles_adjoint_rhs: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_adjoint_rhs: les_model must be 0, 1 or 2");
  if (out == mu || out == y) return fail("sles_adjoint_rhs: out must not alias y or mu");
  if (!sles_supported_n(p->n)) return fail("sles_adjoint_rhs: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  const AdjSrc src{mu, nullptr, 1.0, 0.0, 0.0, 0};
  return p->precision == SLES_FP64 ? adjoint_f64(p, prm, y, src, out, accumulate, linear, dcoeff, cs_grad, work, s)
                                   : adjoint_f32(p, prm, y, src, out, accumulate, linear, dcoeff, cs_grad, work, s);
}

int sles_adjoint_rhs_combined(sles_plan p, const sles_step_params* prm, const void* y, double a, const void* x,
                              double b, const void* z, void* out, double dcoeff, double* cs_grad, void* work,
                              void* stream) {
  if (!p || !prm || !y || !x || !out || !work) return fail("sles_adjoint_rhs_combined: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_adjoint_rhs_combined: les_model must be 0, 1 or 2");
  if (out == x || out == z || out == y) return fail("sles_adjoint_rhs_combined: out must not alias y, x or z");
  if (!sles_supported_n(p->n)) return fail("sles_adjoint_rhs_combined: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  const AdjSrc src{x, z, a, b, 0.0, 0};
  return p->precision == SLES_FP64 ? adjoint_f64(p, prm, y, src, out, 0, 1, dcoeff, cs_grad, work, s)
                                   : adjoint_f32(p, prm, y, src, out, 0, 1, dcoeff, cs_grad, work, s);
}

int sles_etd2_adjoint_rhs(sles_plan p, const sles_step_params* prm, const void* y, const void* lam1,
                          const void* lam2, void* out, double dcoeff, double* cs_grad, void* work, void* stream) {
  if (!p || !prm || !y || !lam1 || !out || !work) return fail("sles_etd2_adjoint_rhs: null argument");
  if (prm->les_model < 0 || prm->les_model > 2) return fail("sles_etd2_adjoint_rhs: les_model must be 0, 1 or 2");
  if (out == lam1 || out == lam2 || out == y) return fail("sles_etd2_adjoint_rhs: out must not alias an input");
  if (!(prm->dt > 0.0)) return fail("sles_etd2_adjoint_rhs: dt must be positive");
  if (!sles_supported_n(p->n)) return fail("sles_etd2_adjoint_rhs: unsupported grid size");
  hipStream_t s = (hipStream_t)stream;
  const AdjSrc src{lam1, lam2, 0.0, 0.0, prm->dt, 1};
  return p->precision == SLES_FP64 ? adjoint_f64(p, prm, y, src, out, 0, 2, dcoeff, cs_grad, work, s)
                                   : adjoint_f32(p, prm, y, src, out, 0, 2, dcoeff, cs_grad, work, s);
}

int sles_etd2_adjoint_combine(sles_plan p, double dt, double nu, const void* lam1, const void* lam2, void* mu,
                              void* lam_out, void* stream) {
  if (!p || !lam1 || !mu || !lam_out) return fail("sles_etd2_adjoint_combine: null argument");
  const long nc = ncplx(p);
  long blocks = (3 * nc + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (p->precision == SLES_FP64)
    hipLaunchKernelGGL(k_etd2_adj<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, nc, p->n, p->nh,
                       dt, nu, (const cpx<double>*)lam1, (const cpx<double>*)lam2, (cpx<double>*)mu,
                       (cpx<double>*)lam_out);
  else
    hipLaunchKernelGGL(k_etd2_adj<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, nc, p->n, p->nh,
                       dt, nu, (const cpx<float>*)lam1, (const cpx<float>*)lam2, (cpx<float>*)mu,
                       (cpx<float>*)lam_out);
  return check_launch("sles_etd2_adjoint_combine");
}

static int axpby_launch(sles_plan p, int nfields, double a, const void* x, double b, const void* y, void* out,
                        void* stream, const char* what) {
  if (!p || !x || !y || !out || nfields < 1) return fail(std::string(what) + ": bad argument");
  const long n = (long)nfields * ncplx(p);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (p->precision == SLES_FP64)
    hipLaunchKernelGGL(k_axpby<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, a,
                       (const cpx<double>*)x, b, (const cpx<double>*)y, (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, a,
                       (const cpx<float>*)x, b, (const cpx<float>*)y, (cpx<float>*)out);
  return check_launch(what);
}

int sles_sum_fields(sles_plan p, int nfields, int nterms, const void* const* terms, void* out, void* stream) {
  if (!p || !terms || !out || nfields < 1) return fail("sles_sum_fields: bad argument");
  if (nterms < 1 || nterms > 8) return fail("sles_sum_fields: nterms must be 1..8");
  SumTerms st{};
  st.n = nterms;
  for (int i = 0; i < nterms; ++i) {
    if (!terms[i]) return fail("sles_sum_fields: null term");
    st.t[i] = terms[i];
  }
  const long n = (long)nfields * ncplx(p);
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (p->precision == SLES_FP64)
    hipLaunchKernelGGL(k_sum_fields<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, st,
                       (cpx<double>*)out);
  else
    hipLaunchKernelGGL(k_sum_fields<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, st,
                       (cpx<float>*)out);
  return check_launch("sles_sum_fields");
}

int sles_axpby(sles_plan p, int nfields, double a, const void* x, double b, void* y, void* stream) {
  return axpby_launch(p, nfields, a, x, b, y, y, stream, "sles_axpby");
}

int sles_lincomb(sles_plan p, int nfields, double a, const void* x, double b, const void* y, void* out,
                 void* stream) {
  return axpby_launch(p, nfields, a, x, b, y, out, stream, "sles_lincomb");
}

}  // extern "C"

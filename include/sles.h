/* libsles: MI355X (gfx950) kernels for the RHS and time-step hot path of the
 * pseudo-spectral Navier-Stokes / LES solver a-k-sangwan/spectral-LES.
 *
 * C ABI.  All array arguments are device pointers to caller-owned memory in
 * the reference's layout (one rank, P = 1):
 *   real field     (N, N, N)        float64 / float32, z fastest
 *   spectral field (N, N, N/2+1)    complex128 / complex64 interleaved, [kx][ky][kz]
 * A "field array" argument (void* const f[k]) lists k such fields.  Every call
 * is asynchronous on the given hipStream_t (passed as void*), returns 0 on
 * success and -1 on error (message from sles_last_error(), per thread), and
 * does no host synchronisation.  Only sles_plan_create / sles_plan_destroy
 * allocate or free device memory.  A plan must not be used from two streams
 * at once.
 *
 * Reference citations are paths inside a-k-sangwan/spectral-LES.
 */
#ifndef SLES_H
#define SLES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLES_ABI_VERSION 1

enum { SLES_FP64 = 0, SLES_FP32 = 1 };
enum { SLES_LES_NONE = 0, SLES_LES_SMAGORINSKY = 1, SLES_LES_VREMAN = 2 };

typedef struct sles_plan_s* sles_plan;

const char* sles_last_error(void);
int sles_version(void);

/* Plan for an N^3 periodic grid in one precision.  Replaces the transform
 * object built by PFFT(comm, [N]*3, axes=[0,1,2], dtype=float, grid=(-1,))
 * at src/initialize_domain.py:59-65.  Supported N: 8 12 16 24 32 48 64 96
 * 128 192 256 384 512. */
int sles_plan_create(sles_plan* plan, int n, int precision);
int sles_plan_destroy(sles_plan plan);
int sles_plan_info(sles_plan plan, int* n, int* nh, int* precision);

/* Kernel variants of the fused step, chosen per plan (by the autotuner,
 * sles/autotune.py, which persists its choices in a JSON table) and of the
 * generic transforms.  Defaults: 0 for the three plan choices, 1 for
 * SLES_TUNE_XCD and SLES_TUNE_FUSE_X; SLES_TUNE_FFT_WIDE 3 at N = 512,
 * 1 (FP64) / 2 (FP32) at N = 256, 0 otherwise; SLES_TUNE_PAD 5 for
 * N >= 128, 0 otherwise.  No counterpart in the reference.
 *   SLES_TUNE_STRIDED_RADIX4  x/y inverse + y forward passes use the radix-4
 *                             line plan (4 points per thread) instead of radix-8
 *   SLES_TUNE_XFWD_RADIX4     same for the x-forward + stage-update pass
 *   SLES_TUNE_Z_RADIX4        same for the fused z kernel
 *   SLES_TUNE_XCD             XCD-aware tile order in the strided passes (fused
 *                             step and the generic transforms)
 *   SLES_TUNE_FUSE_X          RK4 x-forward pass also runs the next stage's
 *                             x-inverse (1, default) or not (0), in both
 *                             precisions at every N; inside one
 *                             sles_step_drive / sles_step_drive_batch call
 *                             the last stage also runs the next step's
 *   SLES_TUNE_FFT_WIDE        generic transforms (sles_fft_r2c / sles_fft_c2r):
 *                             bit 0 = x pass, bit 1 = y pass run tiles of
 *                             twice the columns (256- or 512-byte row runs, LDS tile
 *                             up to 128 KiB); N = 256 and 512 only (ignored
 *                             otherwise); values 0..3
 *   SLES_TUNE_PAD             fused step: the workspace's kx planes (and the
 *                             copies of the RK state that sles_step_drive /
 *                             sles_step_drive_batch step on) start every
 *                             N (N/2+1) values + this many 128-byte lines
 *                             (0..SLES_TUNE_PAD_MAX); 0 steps the caller's
 *                             arrays in place.  sles_step_work_bytes depends
 *                             on it: query it after setting this key
 *   SLES_TUNE_NARROW          fused step, N = 64, 96, 128: the strided and
 *                             x-forward passes take tiles of N (N/2+1) / 1024
 *                             columns (four per CU; 1) instead of / 256 (one
 *                             per CU; 0, default) -- faster for one grid,
 *                             slower for grids that run side by side (batched
 *                             launches, concurrent streams), which fill the
 *                             CUs already; results are identical
 *                             (ignored at other sizes)
 * The "radix-4" alternative plans: 4 points per thread (radix 4) for
 * N = 32, 64, 128, 256, 512; 6 points per thread (radix 3, then 2) for
 * N = 12, 24, 48, 96, 192, 384; ignored for N = 8, 16. */
enum {
  SLES_TUNE_STRIDED_RADIX4 = 0,
  SLES_TUNE_XFWD_RADIX4 = 1,
  SLES_TUNE_Z_RADIX4 = 2,
  SLES_TUNE_XCD = 3,
  SLES_TUNE_FFT_WIDE = 4,
  SLES_TUNE_FUSE_X = 5,
  SLES_TUNE_PAD = 6,
  SLES_TUNE_NARROW = 7
};
#define SLES_TUNE_PAD_MAX 64
int sles_plan_set_tuning(sles_plan plan, int key, int value);
int sles_plan_get_tuning(sles_plan plan, int key, int* value);

/* fft.forward(u, out, normalize=False): batch real fields -> spectral,
 * unscaled, times `scale` (pass 1/N^3 for normalize=True).
 * numpy.fft.rfftn(u, axes=(0,1,2)) semantics.  Calls at
 * src/math_formula.py:10-12,77-125, src/initial.py:13-15. */
int sles_fft_r2c(sles_plan plan, const void* in, void* out, int batch, double scale, void* stream);

/* fft.backward(U, out, normalize=True): batch spectral fields -> real, times
 * `scale` (1/N^3 for normalize=True, 1 otherwise).  numpy.fft.irfftn
 * semantics: x, then y, then z c2r (imaginary parts of the kz = 0 and N/2
 * bins ignored).  `work` holds batch spectral fields; `in` is left unchanged
 * unless `work` == `in`, which is allowed and overwrites it.  Calls at
 * src/math_formula.py:18-24,33-61,
 * src/RK4_method.py:32-34,41-43,97. */
int sles_fft_c2r(sles_plan plan, const void* in, void* out, void* work, int batch, double scale,
                 void* stream);

/* ---- pointwise pieces of src/math_formula.py, smagorinsky.py, vreman.py ---- */

/* c = a x b on real fields (the products inside U_cross_omega,
 * src/math_formula.py:10-12). */
int sles_cross_product(sles_plan plan, void* const a[3], void* const b[3], void* const c[3],
                       void* stream);

/* w = i K x u_hat on spectral fields (the operand of del_cross_U,
 * src/math_formula.py:22-24), wavenumbers of src/initialize_domain.py:76-107. */
int sles_spectral_curl(sles_plan plan, void* const u[3], void* const w[3], void* stream);

/* g[3*i + j] = i K_i u_hat_j (operands of velocity_gradient,
 * src/math_formula.py:33-61; a_{i+1,j+1} = d u_j / d x_i). */
int sles_spectral_gradient(sles_plan plan, void* const u[3], void* const g[9], void* stream);

/* Smagorinsky nu_R = coeff * sqrt(2 S:S), coeff = (Cs dx)^2, from the nine
 * gradient fields a[3*i+j] (src/smagorinsky.py:24-39). */
int sles_smagorinsky_nu(sles_plan plan, void* const a[9], void* nu, double coeff, void* stream);

/* Vreman nu_R = sqrt(B_beta) * [aa > 1e-4 and sqrt(B_beta) > aa/1e7] *
 * sqrt(1/aa) * coeff, coeff = 2.5 Cs^2 dx^2; also writes aa = sum a_ij^2 to
 * `aa` (src/vreman.py:24-123).  NaN for B_beta < 0 as in the reference. */
int sles_vreman_nu(sles_plan plan, void* const a[9], void* nu, void* aa, double coeff, void* stream);

/* tau = (a11, (a12+a21)/2, (a13+a31)/2, a22, (a23+a32)/2, a33) * nu_R: the six
 * operands transformed in residue_stress (src/math_formula.py:73-113). */
int sles_stress_products(sles_plan plan, void* const a[9], const void* nu, void* const tau[6],
                         void* stream);

/* T_i = 2i dealias sum_j K_j tau_hat_ij for tau_hat = F(tau) in the order of
 * sles_stress_products (src/math_formula.py:73-113). */
int sles_stress_divergence(sles_plan plan, void* const tau_hat[6], void* const t[3], void* stream);

/* The spectral tail of computeRHS (src/RK4_method.py:48-70): du *= dealias;
 * P_hat = sum_i (du + T)_i K_i / K^2; du -= P_hat K; du += forcing * u_hat
 * (k != 0); f1 = du (if f1 != NULL); du -= nu K^2 u_hat (if nu > 0);
 * du += T (if t != NULL).  p_hat may be NULL. */
int sles_rhs_finish(sles_plan plan, void* const du[3], void* const t[3], void* const u_hat[3],
                    void* p_hat, void* const f1[3], double nu, double forcing, void* stream);

/* out = K^2 * in for nfields spectral fields (operand of dissipation_rate,
 * src/RK4_method.py:97). */
int sles_mul_k2(sles_plan plan, const void* in, void* out, int nfields, void* stream);

/* x *= s for nfields spectral fields (contiguous). */
int sles_scale(sles_plan plan, void* x, int nfields, double s, void* stream);

/* ---- stage updates (arrays of 3 spectral fields, contiguous) ----
 * kind 0 (RK4, src/RK4_method.py:24-26):
 *    if (c0 != 0) u_hat = u_hat0 + c0 * du ;   u_hat1 += c1 * du
 *    (c0 = b[rk] dt, or 0 for the last stage; c1 = a[rk] dt)
 * kind 1 (RK3 as written, src/RK3_method.py:29): u_hat = u_hat0 + c0 * du
 * kind 2 (src/RK3_method.py:31): u_hat = u_hat0 - c0 * du + (2 c0) * du
 * kind 3 (src/RK3_method.py:33): u_hat = u_hat0 + c0 * (du/6 + 2 du/3 + du/6) */
int sles_stage_update(sles_plan plan, int kind, void* u_hat, const void* u_hat0, void* u_hat1,
                      const void* du, double c0, double c1, void* stream);

/* ETD2 updates on 3 spectral fields, linear operator L = -nu K^2, h = dt.
 * mode 0: the reference formula of src/ETD2_method.py:16-29, literally
 *   u = e^{hL} u + (F1 (1 + h nu K^2) + F0 (1 - h nu K^2 - e^{hL})) / (h (nu K^2)^2)
 *   (denominator 1 at K^2 = 0).
 * mode 1: Cox-Matthews ETD2 (product form)
 *   u = e^{hL} u + h (phi1 + phi2)(hL) F1 - h phi2(hL) F0,
 *   phi functions in FP64 with Taylor series near 0. */
int sles_etd2_update(sles_plan plan, int mode, void* u_hat, const void* f1, const void* f0, double dt,
                     double nu, void* stream);

/* ---- FP64 reductions (results written to device memory) ----
 * `work` must hold sles_reduce_work_bytes() bytes. */
size_t sles_reduce_work_bytes(void);
/* out[0] = sum over nfields real fields of x^2 (U_square numerator,
 * src/RK4_method.py:77-84). */
int sles_sum_squares(sles_plan plan, void* const f[], int nfields, double* out, void* work,
                     void* stream);
/* out[0] = sum_f sum_x a_f b_f (dissipation_rate sum, src/RK4_method.py:94-98). */
int sles_dot(sles_plan plan, void* const a[], void* const b[], int nfields, double* out, void* work,
             void* stream);
/* out[0] = signed max over nfields real fields (numpy.max of the list U,
 * src/SM_solver.py:245-255). */
int sles_max(sles_plan plan, void* const f[], int nfields, double* out, void* work, void* stream);
/* Binned spectrum of calculate_Ek (src/SM_solver.py:154-192): for bins
 * b = 0..nbins-1 with K^2 in [(5b)^2, (5b+5)^2):
 *   ek_sum[b] = sum Ek, ksum[b] = sum sqrt(K^2), count[b] = number of modes,
 * Ek = sum_i |u_hat_i|^2, doubled for kz >= 1.  u_hat = 3 contiguous fields.
 * Deterministic: the same input gives the same bits on every call (fixed
 * summation order, no atomics).  nbins <= 640 (every mode of N = 512 falls in
 * 91 bins).  `work` holds sles_spectrum_work_bytes(plan, nbins) bytes. */
size_t sles_spectrum_work_bytes(sles_plan plan, int nbins);
int sles_spectrum(sles_plan plan, const void* u_hat, int nbins, double* ek_sum, double* ksum,
                  double* count, void* work, void* stream);

/* Spectrum over unit shells, the E(k) of a spectral loss:
 * e[k-1] = sum over modes with |K| in [k - 1/2, k + 1/2) of
 * (1/2) w sum_i |u_hat_i|^2 / N^6, k = 1..kmax, w = 1 on the kz = 0 and
 * kz = N/2 planes and 2 elsewhere, kmax <= 640 (N = 512 has 443 shells).
 * Deterministic as sles_spectrum; `work` as for sles_spectrum with
 * nbins = kmax.  sles_shell_spectrum_vjp writes the
 * gradient of sum_k grad_e[k-1] e[k-1] with respect to u_hat (as
 * d/dRe + i d/dIm, 3 fields like u_hat): grad_e[k-1] w u_hat / N^6 in shells
 * 1..kmax, 0 elsewhere.  No counterpart in the reference (the loss of
 * differentiable runs). */
int sles_shell_spectrum(sles_plan plan, const void* u_hat, int kmax, double* e, void* work, void* stream);
int sles_shell_spectrum_vjp(sles_plan plan, const void* u_hat, int kmax, const double* grad_e, void* grad_u,
                            void* stream);

/* ---- fused time step (src/RK4_method.py:20-71, src/RK3_method.py:20-41,
 * src/ETD2_method.py:7-37 with src/smagorinsky.py / src/vreman.py) ----
 * One call advances u_hat by one step of `scheme` with every RHS evaluation
 * done in five transform passes (x, y inverse; a z kernel that does the c2r,
 * all physical-space work and the r2c; y, x forward with dealiasing,
 * projection, forcing, viscous term and the stage update fused in).
 *
 *   SLES_SCHEME_RK4   classical RK4 as in RK4_method.solveForNextTimestep;
 *                     needs u_hat1 == u_hat on entry (SM_solver.py:33) and
 *                     leaves u_hat0, u_hat1 as the reference does.  If aux is
 *                     not NULL it receives F(u_hat) at the start state (the
 *                     nonlinear term below), the F_{n-1} of a following ETD2.
 *   SLES_SCHEME_RK3   the reference RK3 as written (first order).
 *   SLES_SCHEME_ETD2  Cox-Matthews ETD2: aux holds F_{n-1} on entry and F_n
 *                     on exit; F includes the SGS stress divergence.
 *   SLES_SCHEME_RHS   aux = computeRHS(u_hat) (full right-hand side).
 *   SLES_SCHEME_NONLINEAR aux = F(u_hat) = computeRHS without -nu K^2 u_hat.
 *
 * dt_device = 1: dt = min(cfl dx / max(u), 0.1 dx^2 / nu) from the signed
 * max of u at the start state (find_max_velocity/compute_dt,
 * src/SM_solver.py:245-277), computed on the device into scalars[0] (max u)
 * and scalars[1] (dt); when that dt is NaN or below 1e-6 the step is not
 * taken (u_hat unchanged) -- the host checks scalars[1].
 * u_out / curl_out (3 real fields each, may be NULL) receive U and curl of
 * the new state, as the reference leaves them.  `work` must hold
 * sles_step_work_bytes(plan, les_model) bytes. */
enum {
  SLES_SCHEME_RK4 = 0,
  SLES_SCHEME_RK3 = 1,
  SLES_SCHEME_ETD2 = 2,
  SLES_SCHEME_RHS = 3,
  SLES_SCHEME_NONLINEAR = 4
};

typedef struct {
  double dt;        /* step size when dt_device == 0 */
  double nu;        /* kinematic viscosity */
  double forcing;   /* linear forcing coefficient f (f * u_hat for k != 0) */
  double sgs_coeff; /* (Cs dx)^2 Smagorinsky, 2.5 Cs^2 dx^2 Vreman */
  double cfl;       /* CFL number for dt_device */
  double dx;        /* 2 pi / N */
  int les_model;    /* SLES_LES_NONE / SMAGORINSKY / VREMAN */
  int dt_device;    /* 0: use dt; 1: CFL dt on the device */
} sles_step_params;

size_t sles_step_work_bytes(sles_plan plan, int les_model);

/* The driver loop of spectralMethodSolver (src/SM_solver.py:18-68) on the
 * device: nsteps fused steps of `scheme` (RK4, RK3 or ETD2; params->dt and
 * dt_device are not used) with dt and the loop's stop rules evaluated per
 * step on the device, so the host enqueues a chunk of steps with no
 * synchronisation and reads `state` once afterwards.  Each step:
 *   1. max(u) of the state (signed, as numpy.max of U) and
 *      dt = min(CFL dx / max u, 0.1 dx^2 / nu), CFL = cfl_first until the
 *      first step is taken (0.3 in the reference, :21-23), params->cfl after
 *      (:60-63);
 *   2. dt NaN: status NAN (the loop's break, :64-67); dt < 1e-6: status
 *      DT_TOO_SMALL (compute_dt's exit, :273-276);
 *   3. t = dt if state[T] is NaN (t starts at the first dt, :25);
 *   4. loop top: t >= state[T_PAUSE] stops with status PAUSED so that the
 *      host can write the outputs due at this t (:40-45), unless the status
 *      on entry is RESUME (the host has written them); t >= state[T_FINAL]
 *      or dt < 1e-8: status DONE (:46-47);
 *   5. t += dt (:51), steps += 1, and the step is taken with dt.
 * Once the status is not RUNNING (or RESUME) the remaining steps of the
 * call leave U_hat unchanged.  The caller starts a run with t = NaN,
 * status = RUNNING, steps = 0 and sets T_PAUSE (+inf for none) and T_FINAL;
 * state holds SLES_DRIVE_STATE_SIZE doubles in device memory.  u_out /
 * curl_out receive U and curl of the state at the end of the call (the
 * state the reference leaves); they are written by the call's last step only,
 * since no step reads them (each recomputes its physical fields from U_hat)
 * and nothing else runs between the steps of a call.  `work` as for
 * sles_step; with SLES_TUNE_PAD > 0 the call copies U_hat (and ETD2's aux)
 * into the workspace's padded layout, steps that copy and writes U_hat,
 * U_hat0, U_hat1 and aux back on return (U_hat0, U_hat1 and RK4's aux only
 * if a step was taken), so the arrays hold the reference's values between
 * calls.  No counterpart in the
 * reference as an entry point (its loop is host Python). */
enum {
  SLES_DRIVE_STATE_UMAX = 0,     /* max u of the last state examined */
  SLES_DRIVE_STATE_DT = 1,       /* dt from the CFL rule of that state */
  SLES_DRIVE_STATE_T = 2,        /* the loop's t (NaN before the run) */
  SLES_DRIVE_STATE_STATUS = 3,   /* SLES_DRIVE_* below */
  SLES_DRIVE_STATE_STEPS = 4,    /* steps taken in this run */
  SLES_DRIVE_STATE_DT_APPLY = 5, /* dt of the last step, NaN if not taken */
  SLES_DRIVE_STATE_T_PAUSE = 6,  /* next output time (set by the host) */
  SLES_DRIVE_STATE_T_FINAL = 7,  /* final time T (set by the host) */
  SLES_DRIVE_STATE_SIZE = 8
};
enum {
  SLES_DRIVE_RUNNING = 0,
  SLES_DRIVE_DONE = 1,
  SLES_DRIVE_NAN = 2,
  SLES_DRIVE_DT_TOO_SMALL = 3,
  SLES_DRIVE_PAUSED = 4,
  SLES_DRIVE_RESUME = 5
};
int sles_step_drive(sles_plan plan, int scheme, const sles_step_params* params, double cfl_first, int nsteps,
                    void* u_hat, void* u_hat0, void* u_hat1, void* aux, void* u_out, void* curl_out, double* state,
                    void* work, void* stream);

/* sles_step_drive for `members` independent runs of one grid size (an
 * ensemble of seeds and SGS coefficients), advanced by the same launches:
 * every pass of a step is one launch over all members (blockIdx.y = member),
 * so small grids fill the device together.  Each member is one run of
 * spectralMethodSolver's loop (src/SM_solver.py:18-68) with its own state
 * (own dt, t, status and stop rules, exactly as sles_step_drive); the
 * members share the plan, the scheme and every field of `params` except
 * sgs_coeff, which comes per member from the device array sgs_coeff
 * (`members` doubles; may be NULL without LES).  Memory is packed by member:
 * member m's u_hat (and u_hat0, u_hat1, aux) starts 3 C elements after member
 * m-1's (C = N * N * (N/2+1)), its driver state SLES_DRIVE_STATE_SIZE doubles
 * after; `work` holds sles_step_batch_work_bytes(plan, les_model, members)
 * bytes.  No U / curl output (read a member's U_hat instead).  No counterpart
 * in the reference, which runs one field per process. */
#define SLES_BATCH_MAX_MEMBERS 65535
size_t sles_step_batch_work_bytes(sles_plan plan, int les_model, int members);
int sles_step_drive_batch(sles_plan plan, int scheme, const sles_step_params* params, int members,
                          const double* sgs_coeff, double cfl_first, int nsteps, void* u_hat, void* u_hat0,
                          void* u_hat1, void* aux, double* state, void* work, void* stream);

/* sles_step without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a static device array of SLES_STEP_SCRATCH_BYTES (64 MiB)
 * that every device running the library holds until the process ends; its
 * address is taken per device by sles_plan_create.  Calls whose workspace
 * -- the W fields and CFL partials of sles_step_work_bytes, without the
 * drivers' copies of the RK state that it also counts when SLES_TUNE_PAD > 0
 * -- exceeds it are rejected with an error: with LES up to N = 96 in FP64
 * and N = 96 in FP32 (128 in FP32 needs 73 MiB), without LES up to 96
 * (FP64) and 128 (FP32).  Every call on a device shares that
 * device's array, so calls that overlap on one device (two streams, two
 * threads) race; each call writes the array before it reads it, so nothing
 * carries over between calls.  Same arguments, results and errors as
 * sles_step otherwise; sles_profile_enable does not time it. */
#define SLES_STEP_SCRATCH_BYTES (64u << 20)
int sles_step_noalloc(sles_plan plan, int scheme, const sles_step_params* params, void* u_hat, void* u_hat0,
                      void* u_hat1, void* aux, void* u_out, void* curl_out, double* scalars, void* stream);

/* Kernel timing of sles_step for measurement: while enabled, every launch of
 * sles_step is bracketed by hipEvents on its stream.  sles_profile_read
 * synchronises on them, adds the elapsed milliseconds and launch counts per
 * kernel kind (0 x-inverse, 1 y-inverse, 2 z, 3 y-forward, 4 x-forward with
 * the stage update, 5 CFL reduction, 6 U/curl output) into ms[]/count[]
 * (nkinds entries) and clears the record.  The profiling state is global to
 * the process and not thread-safe; do not enable it under HIP graph capture. */
int sles_profile_enable(int on);
int sles_profile_read(double* ms, long* count, int nkinds);
int sles_step(sles_plan plan, int scheme, const sles_step_params* params, void* u_hat, void* u_hat0, void* u_hat1,
              void* aux, void* u_out, void* curl_out, double* scalars, void* work, void* stream);

/* One fused RK4 step (as sles_step with SLES_SCHEME_RK4, host dt, no physical
 * output) that also writes the stage states Y_1, Y_2, Y_3 (the arguments of
 * the second to fourth RHS evaluations, 3 spectral fields each) to stages
 * (9 fields).  u_hat0 receives a copy of u_hat (Y_0); u_hat1 is written at
 * the first stage before it is read, so its content on entry is not used.
 * No counterpart in the reference: the adjoint's recomputation of a step. */
int sles_rk4_stage_states(sles_plan plan, const sles_step_params* params, void* u_hat, void* u_hat0, void* u_hat1,
                          void* stages, void* work, void* stream);

/* ---- hand-written adjoint (no counterpart in the reference) ----
 * out (=, or += when accumulate) J(y)^T mu: the vector-Jacobian product of
 * computeRHS (src/RK4_method.py:38-71, with the closure of les_model) at the
 * state y, for an adjoint mu of the RHS; complex arrays are taken as real
 * vectors (Re, Im), the convention of torch gradients.  linear = 1 includes
 * the linear part (forcing [k != 0] - nu K^2); linear = 2 only the forcing
 * (the nonlinear term F that ETD2 uses); linear = 0 neither.  If
 * cs_grad is not NULL, cs_grad[0] += sum_x nu_bar d nu_R / d Cs with
 * d sgs_coeff / d Cs = dcoeff (2 Cs dx^2 Smagorinsky, 5 Cs dx^2 Vreman).
 * `work` must hold sles_adjoint_work_bytes(plan, les_model) bytes. */
size_t sles_adjoint_work_bytes(sles_plan plan, int les_model);
int sles_adjoint_rhs(sles_plan plan, const sles_step_params* params, const void* y, const void* mu, void* out,
                     int accumulate, int linear, double dcoeff, double* cs_grad, void* work, void* stream);
/* The RK4 adjoint stage in one call: sles_adjoint_rhs with linear = 1 and
 * accumulate = 0 for
 *   mu = a x + b z   (z may be NULL, taken as 0)
 * formed inside the adjoint passes (no separate combination pass).  out must
 * not alias y, x or z; `work` as for sles_adjoint_rhs. */
int sles_adjoint_rhs_combined(sles_plan plan, const sles_step_params* params, const void* y, double a, const void* x,
                              double b, const void* z, void* out, double dcoeff, double* cs_grad, void* work,
                              void* stream);
/* ETD2 adjoint combination (Cox-Matthews, L = -nu K^2, h = dt, FP64 phi):
 *   mu = h (phi1 + phi2)(hL) lam1 - h phi2(hL) lam2 ;  lam_out = e^{hL} lam1
 * lam2 may be NULL (taken as 0).  3 spectral fields each. */
int sles_etd2_adjoint_combine(sles_plan plan, double dt, double nu, const void* lam1, const void* lam2, void* mu,
                              void* lam_out, void* stream);
/* One ETD2 backward step in one call (the combination above fused into the
 * adjoint passes, mu never stored):
 *   out = e^{hL} lam1 + J_F(y)^T (h (phi1 + phi2)(hL) lam1 - h phi2(hL) lam2)
 * with h = params->dt, J_F the Jacobian of the nonlinear term (linear = 2
 * above: forcing, no viscous term) and the cs_grad accumulation of
 * sles_adjoint_rhs.  lam2 may be NULL (taken as 0); out must not alias the
 * inputs; `work` as for sles_adjoint_rhs. */
int sles_etd2_adjoint_rhs(sles_plan plan, const sles_step_params* params, const void* y, const void* lam1,
                          const void* lam2, void* out, double dcoeff, double* cs_grad, void* work, void* stream);
/* out = terms[0] + terms[1] + ... (in that order) over nfields spectral
 * fields, 1 <= nterms <= 8 (terms: host array of device pointers; out may be
 * one of the terms). */
int sles_sum_fields(sles_plan plan, int nfields, int nterms, const void* const* terms, void* out, void* stream);
/* y = a x + b y over nfields spectral fields (contiguous). */
int sles_axpby(sles_plan plan, int nfields, double a, const void* x, double b, void* y, void* stream);
/* out = a x + b y over nfields spectral fields (out may be x or y). */
int sles_lincomb(sles_plan plan, int nfields, double a, const void* x, double b, const void* y, void* out,
                 void* stream);

/* ---- measurement (no counterpart in the reference) ----
 * The loads and stores of a strided x pass with no arithmetic: every thread
 * of a tile of consecutive columns reads its rows of the nin spectral
 * fields at `in` and writes their sum to the nout fields at `out`; fields
 * in the fused step's workspace layout (kx planes N (N/2+1) values +
 * SLES_TUNE_PAD 128-byte lines apart, fields N planes apart).  Its
 * bandwidth is the ceiling of a pass with that tile shape:
 * tile 0 the radix-4 x-forward tile (FP64 8 columns x 64 threads, 128-byte
 * row runs; FP32 16 x 64), 1 16-column tiles (the radix-8 strided shape),
 * 2 32-column tiles.
 * N = 256 and 512 only; honours SLES_TUNE_XCD. */
int sles_stream_probe(sles_plan plan, const void* in, int nin, void* out, int nout, int tile, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SLES_H */

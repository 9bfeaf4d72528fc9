/*
 * libhydft — MI355X (gfx950) HIP implementation of HyDFT's 3-D pseudo-spectral
 * right-hand side and low-storage RK3 time loop, behind a plain C ABI.
 *
 * Conventions (those of the reference, abdoudiaw/HyDFT):
 *  - Real fields are (nx, ny, nz) arrays with x fastest (Fortran column-major),
 *    index i + nx*(j + ny*k).  Half spectra are (nkx = nx/2+1, ny, nz) complex
 *    (interleaved re, im doubles), x fastest: FFTW's r2c_3d(nz, ny, nx) layout.
 *  - Forward transforms are normalised by 1/(nx*ny*nz); backward are not and
 *    preserve their input (src/core/fft.f90:52-69).
 *  - Vector fields are component-major: component c at offset c*nx*ny*nz.
 *  - Stress components are (xx, yy, zz, xy, xz, yz) (src/closure/stress.f90:11).
 *  - Every call returns HY_OK (0) or an error code; hy_last_error() gives the
 *    message.  The Fortran binding maps a nonzero code to fatal() (error stop 1),
 *    as the reference does (src/core/kinds.f90:26-30).
 *  - Device-pointer entry points (suffix-free `d_` arguments) are stream
 *    ordered on `stream` (a hipStream_t, NULL = default stream).  Host-pointer
 *    entry points (`h_` arguments) synchronise before returning.
 *  - Handles are not thread-safe; one model per host thread at a time.
 */
#ifndef HYDFT_H
#define HYDFT_H

#ifdef __cplusplus
extern "C" {
#endif

#define HY_OK 0
#define HY_ERR_INVALID 1        /* invalid argument or grid (grid.f90:45-48)            */
#define HY_ERR_NONPOSITIVE 2    /* density became non-positive (model.f90:91)          */
#define HY_ERR_OOM 3            /* device allocation failed                            */
#define HY_ERR_HIP 4            /* HIP runtime error                                   */
#define HY_ERR_UNKNOWN_KIND 5   /* unknown term / closure / potential name             */
#define HY_ERR_UNSUPPORTED 6    /* grid size outside {4..256} of the form 2^a 3^b      */
#define HY_ERR_NO_DT 7          /* cannot determine a time step (model.f90:180)        */
#define HY_ERR_BUSY 8           /* serving: no slot or HBM budget for the request      */

#define HY_PREC_FP64 0          /* everything in FP64                                  */
#define HY_PREC_MIXED 1         /* FP32 transforms and functional kernels, FP64 state  */

typedef struct hy_grid_s* hy_grid;
typedef struct hy_model_s* hy_model;

/* ---- library ----------------------------------------------------------- */
int hy_init(int device);
const char* hy_last_error(void);
int hy_version(void);
int hy_device_sync(void);

/* ---- device memory for callers without a device allocator (the Fortran host) */
int hy_device_alloc(unsigned long long bytes, void** d_out);
int hy_device_free(void* d);
int hy_copy_to_device(void* d_dst, const void* h_src, unsigned long long bytes);
int hy_copy_to_host(void* h_dst, const void* d_src, unsigned long long bytes);

/* ---- grid_t + fft_t (src/core/grid.f90:34-113, src/core/fft.f90:36-49) -- */
int hy_grid_create(int nx, int ny, int nz, double lx, double ly, double lz, int dealias,
                   int precision, hy_grid* out);
int hy_grid_destroy(hy_grid g);
/* nkx, ndim, kept box (Kx, Ky, Kz), volume, kmax_dealiased (grid.f90:15-30) */
int hy_grid_query(hy_grid g, int* nkx, int* ndim, int* kept3, double* volume,
                  double* kmax_dealiased);
/* the 2/3-rule mask on the half spectrum, (nkx, ny, nz) doubles (grid.f90:89-112) */
int hy_grid_mask(hy_grid g, double* h_mask);

/* ---- fft_t (src/core/fft.f90:52-69), FP64 device buffers ---------------- */
int hy_fft_r2c(hy_grid g, const double* d_u, double* d_uhat, void* stream);
int hy_fft_c2r(hy_grid g, const double* d_uhat, double* d_u, void* stream);

/* ---- pencil plans ---------------------------------------------------------
   The y/z transform passes run with a column count per workgroup chosen by
   timing the candidates once per pass shape (precision, length, direction,
   column count, outer count, fields, batch size) on its first run outside a
   graph capture or a batch recording, and kept for the process: the role of
   FFTW_MEASURE in fft_t%init (fft.f90:46-47).  Results do not depend on the
   plan.  hy_plan_tuning(0) keeps the untimed defaults (as HYDFT_TUNE=0 does);
   hy_plan_reset forgets the choices; hy_plan_list writes up to `max` entries
   {fp64, L, inverse, A, B, nfields, batch, columns per workgroup} and returns
   the number of entries. */
int hy_plan_tuning(int on);
int hy_plan_reset(void);
int hy_plan_list(int max, int* out8);

/* ---- spec_t (src/core/spectral.f90:67-188), FP64 device buffers --------- */
int hy_spec_grad_hat(hy_grid g, const double* d_uhat, int idir, double* d_du, void* stream);
int hy_spec_grad(hy_grid g, const double* d_u, double* d_du3, void* stream);
int hy_spec_lap(hy_grid g, const double* d_u, double* d_lu, void* stream);
int hy_spec_div(hy_grid g, const double* d_v3, double* d_dv, void* stream);
int hy_spec_convolve(hy_grid g, const double* d_kernel, const double* d_uhat, double* d_cu,
                     void* stream);
int hy_spec_dealias(hy_grid g, double* d_u, void* stream);
int hy_spec_mean(hy_grid g, const double* d_u, double* h_mean, void* stream);
int hy_spec_mode(hy_grid g, const double* d_u, int mx, int my, int mz, double* h_re_im,
                 void* stream);
/* hy_spec_mean without device allocation.  No counterpart in the reference:
   for callers that cannot allocate device memory (tight serving loops,
   harnesses that forbid hipMalloc).  hy_spec_mean takes its partial sums from
   the grid's operator scratch (eight half spectra, allocated on a grid's first
   operator call); this variant keeps them in a fixed 8 KiB device array per
   GPU (513 doubles used, for every grid size) and calls no allocator.  Same
   kernels, same result bit for bit.  Calls that overlap on one GPU share the
   array and race: serialise them. */
int hy_spec_mean_noalloc(hy_grid g, const double* d_u, double* h_mean, void* stream);

/* ---- fermi_dirac.f90 functions (host): which = 0 I_{-1/2}, 1 I_{1/2},
   2 I_{3/2}, 3 I'_{-1/2}, 4 I''_{-1/2}, 5 xi, 6 dxi/dalpha, 7 inverse of I_{1/2}
   (fermi_dirac.f90:150-267); 8-12 the forms the field kernels evaluate,
   computed on the host from the same source: 8 I_{-1/2}, 9 I'_{-1/2},
   10 I''_{-1/2}, 11 I_{1/2}, 12 inverse of I_{1/2} */
int hy_fd_eval(int which, int n, const double* h_x, double* h_y);

/* ---- stress_t (src/closure/stress.f90:60-105), code-unit eta, xi -------- */
int hy_stress_pi0(hy_grid g, double eta, double xi, const double* d_u3, double* d_pi6,
                  void* stream);
int hy_stress_div(hy_grid g, const double* d_pi6, double* d_f3, void* stream);

/* ---- free-energy terms (src/functional/*.f90) ---------------------------- */
#define HY_TERM_IDEAL 0     /* fe_ideal_classical.f90 */
#define HY_TERM_TF 1        /* fe_kinetic_tfk.f90 fe_tf_t */
#define HY_TERM_GRADIENT 2  /* fe_kinetic_tfk.f90 fe_gradient_t */
#define HY_TERM_HARTREE 3   /* fe_hartree.f90 */
#define HY_TERM_RY 4        /* fe_correlation_ry.f90 */

typedef struct {
  int kind;                 /* HY_TERM_*                                       */
  double n0, beta;          /* fe_term_t%n0, %beta (free_energy.f90:23-25)     */
  double e2, kappa;         /* hartree / ry                                    */
  double cn, alpha0, cg;    /* tf / gradient (units.f90:31, species.f90:127)   */
  int subtract_mean_field;  /* ry (species.f90:136)                            */
  int ntab;                 /* ry: radial table size                           */
  const double* k_tab;      /* ry: host table k (increasing, > 0)              */
  const double* c_tab;      /* ry: host table c(k) (structure_source.f90:18-26)*/
} hy_term_desc;

/* ---- functional_t and its fe_term_t terms on FP64 device fields ---------
   The reference's physics plug-in interface (free_energy.f90:22-59,
   functional.f90:17-90) on a drop-in grid g, which must outlive the handle:
   the terms' radial tables are copied at create.  d_n is a real field, d_nhat
   its forward transform (hy_fft_r2c layout), d_mu a real field; stream ordered
   on `stream` and synchronised at return (the density check).  A term index
   is the position in `terms`; -1 in hy_fun_kernel / hy_fun_energy means the
   whole functional.  HY_ERR_NONPOSITIVE where a term needs n > 0
   (fe_ideal_classical.f90:39, fe_kinetic_tfk.f90:68,122); mu is then
   unspecified. */
typedef struct hy_fun_s* hy_fun;
/* functional_add + functional_setup (functional.f90:32-53) */
int hy_fun_create(hy_grid g, int nterms, const hy_term_desc* terms, hy_fun* out);
int hy_fun_destroy(hy_fun f);
int hy_fun_nterms(hy_fun f, int* n);
/* fe_term_t%add_mu (free_energy.f90:39-46): mu += dF_term/dn */
int hy_fun_add_mu(hy_fun f, int term, const double* d_n, const double* d_nhat, double* d_mu, void* stream);
/* functional_mu (functional.f90:56-67): mu = 0, then every term's add_mu in list order */
int hy_fun_mu(hy_fun f, const double* d_n, const double* d_nhat, double* d_mu, void* stream);
/* fe_term_t%kernel (free_energy.f90:47-51), functional_kernel (functional.f90:70-78) */
int hy_fun_kernel(hy_fun f, int term, double k, double* h_K);
/* fe_term_t%energy (free_energy.f90:52-58), functional_energy (functional.f90:80-90) */
int hy_fun_energy(hy_fun f, int term, const double* d_n, const double* d_nhat, double* h_F, void* stream);

/* ---- model_t (src/model/model.f90) -------------------------------------- */
#define HY_CLOSURE_NEWTONIAN 0
#define HY_CLOSURE_MAXWELL 1
#define HY_VEXT_NONE 0
#define HY_VEXT_MODE 1
#define HY_VEXT_DRIVEN 2
#define HY_VEXT_GAUSSIAN 3

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz;
  int dealias;              /* &grid dealias (params.f90:21)                    */
  int precision;            /* HY_PREC_*                                        */
  double mass, n0;
  int nterms;
  const hy_term_desc* terms;
  int closure;              /* HY_CLOSURE_* (stress.f90:43-52)                  */
  double eta, xi, tau;      /* stress_t%eta, %xi in code units, tau (stress.f90:53-54) */
  double eta_l;             /* transport_t%eta_l for compute_dt (model.f90:165,170) */
  int vext_kind;            /* HY_VEXT_* (external_potential.f90:49-58)         */
  double vext_amp;
  int vext_mode[3];
  double vext_omega, vext_width, vext_ramp;
  double dt, cfl;           /* &time dt (> 0 fixed) and cfl (params.f90:45-46)   */
} hy_model_desc;

int hy_model_create(const hy_model_desc* desc, hy_model* out);
int hy_model_destroy(hy_model m);
/* state in the reference layout; projected onto the kept band on entry.
   A new state clears the non-positive-density and no-time-step conditions
   of the previous one; hy_model_set_dt clears the no-time-step condition. */
int hy_model_set_state(hy_model m, const double* h_n, const double* h_mom3, const double* h_pi6);
int hy_model_get_state(hy_model m, double* h_n, double* h_mom3, double* h_pi6);
int hy_model_set_time(hy_model m, double t, int step);
int hy_model_get_time(hy_model m, double* t, double* dt, int* step);
int hy_model_set_dt(hy_model m, double dt);
/* model_compute_dt (model.f90:150-183) */
int hy_model_compute_dt(hy_model m, double* h_dt);
/* model_advance x nsteps (model.f90:186-227).  dt_every = 0 keeps the current
   dt; k > 0 calls compute_dt before steps 0, k, 2k, ... of this call (the
   driver cadence of src/hydft.f90:42-43 with k = 20).  tend > 0 clamps the
   last step to tend and stops there (src/hydft.f90:41,44); tend < 0 stops
   once t >= -tend without clamping (the loops of the reference's tests).  When probe3 is not
   NULL, mode probe3 of n/n0 and mom_x (diagnostics.f90:70-77) is recorded
   after every step into h_probe (6 doubles per step: t, Re, Im dn_k/n0,
   Re, Im mom_k, dt).  Returns the number of steps taken in *h_taken. */
int hy_model_advance(hy_model m, int nsteps, int dt_every, double tend, const int* probe3,
                     double* h_probe, int* h_taken);
/* model_rhs (model.f90:85-146) at time t from the current state */
int hy_model_rhs(hy_model m, double t, double* h_dn, double* h_dmom3, double* h_dpi6);
/* functional_mu (functional.f90:56-67) of a given real density */
int hy_model_mu(hy_model m, const double* h_n, double* h_mu);
/* diag_budgets (diagnostics.f90:41-60): mass, mom x/y/z, E_kin, F */
int hy_model_budgets(hy_model m, double* h_out6);
/* spec_mode of n/n0 and mom_x at (mx, my, mz) (diagnostics.f90:70-77) */
int hy_model_probe(hy_model m, int mx, int my, int mz, double* h_out4);
/* functional_kernel K(k) (functional.f90:64-72) */
int hy_model_kernel(hy_model m, double k, double* h_K);

/* ---- measurement (no counterpart in the reference) ------------------------
   hy_model_profile(m, 1) clears and starts recording a HIP event pair around
   every kernel launch of the model, on the model's stream; _read returns, per
   kernel class, the launch count, the summed event time (ms) and the summed
   algorithmic bytes (compulsory reads + writes of each launch), and returns
   the number of classes.  Class names come from hy_profile_class_name.
   hy_model_profile_classes(m, mask) records only the classes c whose bit
   (1 << c) is set in mask (0: off): each event record costs a few us of
   stream time, so a timed run brackets only the class it measures. */
int hy_model_profile(hy_model m, int enable);
int hy_model_profile_classes(hy_model m, unsigned mask);
int hy_model_profile_read(hy_model m, int maxcls, int* launches, double* total_ms, double* bytes);
const char* hy_profile_class_name(int cls);

/* ---- batched and served runs: captured steps, stream state -------------
   No counterpart in the reference (it runs one model at a time on the CPU).
   hy_model_graph_build captures nsteps RK3 steps of the model as
   hy_model_advance(m, nsteps, dt_every, 0, probe3, ...) takes them (dt
   recomputed every dt_every steps, nsteps a multiple of it; no tend clamp)
   into a HIP graph without advancing the model.  With probe3, every step
   writes its probe row (6 doubles, as hy_model_advance) to slot
   (step - 1) % ring of a device ring of `ring` >= nsteps rows, step being the
   model's step count after the step.  hy_model_graph_launch enqueues replays
   on the model's stream and returns at once; each replay advances the model by
   nsteps.  hy_model_graph_probe synchronises and copies the ring (ring rows).
   hy_model_idle reports whether the model's queued work has finished;
   hy_model_check synchronises and returns the non-positive-density and
   no-time-step conditions of the work so far.  hy_model_device_bytes is the
   model's device memory (admission by HBM budget). */
int hy_model_graph_build(hy_model m, int nsteps, int dt_every, const int* probe3, int ring);
int hy_model_graph_launch(hy_model m, int nlaunch);
int hy_model_graph_probe(hy_model m, double* h_probe);
int hy_model_idle(hy_model m, int* idle);
int hy_model_check(hy_model m);
int hy_model_device_bytes(hy_model m, unsigned long long* bytes);

/* ---- batches of same-shape models: one launch per kernel for all ---------
   No counterpart in the reference.  For many small runs the device's launch
   rate, not its bandwidth, bounds a step (a step is ~45 launches); a batch
   launches each kernel of the step once for all its models (grid z = model).
   hy_batch_create records nsteps steps of each model as hy_model_graph_build
   would capture them (same arguments and preconditions) and checks that every
   model yields the same kernels and grids (same grid, precision, closure and
   options; each model at most once); the models' own state, tables and probe
   rings are used in place.
   hy_batch_launch enqueues nlaunch passes of those nsteps steps on the
   batch's stream and returns; hy_batch_idle / hy_batch_sync poll / wait.
   While a batch is in flight its models must not be used otherwise; a model
   stays usable after hy_batch_destroy (its step count and time then read the
   device's).  hy_model_probe_ring copies `ring` probe rows of a model. */
typedef struct hy_batch_s* hy_batch;
int hy_batch_create(const hy_model* models, int n, int nsteps, int dt_every, const int* probe3, int ring,
                    hy_batch* out);
int hy_batch_launch(hy_batch b, int nlaunch);
/* Runs of different lengths in one batch: model i takes no further step once
   its step count reaches stops[i] (0: no limit); every kernel of its later
   launches returns at once, so its state, time and probe ring stay those after
   step stops[i], exactly as when it runs alone.  A stop must be a multiple of
   the batch's nsteps (steps per launch), counted from step 0; the limits are
   stream ordered after the launches already queued and are lifted when the
   batch is destroyed. */
int hy_batch_set_stops(hy_batch b, const int* stops);
int hy_batch_idle(hy_batch b, int* idle);
/* Groups of batches: runs on uneven grids advanced by one launch.
   hy_group_create captures one pass of each batch's launches into one HIP
   graph, each batch on its own branch (the branches are independent, so the
   device overlaps them); hy_group_launch enqueues nlaunch replays on the
   group's stream and returns: every run of every batch then advances by its
   batch's nsteps per replay (step limits of hy_batch_set_stops hold).  While a
   group is in flight its batches and models must not be used otherwise;
   destroy the group before its batches.  No counterpart in the reference. */
typedef struct hy_group_s* hy_group;
int hy_group_create(const hy_batch* batches, int n, hy_group* out);
int hy_group_launch(hy_group g, int nlaunch);
int hy_group_idle(hy_group g, int* idle);
int hy_group_sync(hy_group g);
int hy_group_destroy(hy_group g);
int hy_batch_sync(hy_batch b);
int hy_batch_destroy(hy_batch b);
int hy_model_probe_ring(hy_model m, int ring, double* h_probe);
/* The kept box (Kx, Ky, Kz) of the model's grid (grid.f90:90-108).  Models
   of the same nx, ny, nz can differ in it: when N is divisible by 3 the
   boundary mode sits exactly on the 2/3-rule threshold and the comparison
   depends on the box length's rounding; only models of the same kept box
   share a batch. */
int hy_model_kept(hy_model m, int* kept3);

/* ---- batched OZ-HNC direct solve (oz_hnc.f90:111-140 direct branch, the
   iteration of oz_hnc.f90:142-296) --------------------------------------
   nsolve independent solves on one radial grid r_j = (j+1) dr, k_m = (m+1) dk
   (dk = pi / ((nr+1) dr)), iterated on the GPU: per solve, n0, beta v_short(r)
   (h_bvs, nsolve x nr), beta v_long(k) (h_bvlk) and the initial gamma_s
   (h_gs0, the mean-field guess or a continuation start).  Picard + Ng
   iteration with the reference's rules (Ng after 5 iterations with three
   history entries, cooldown of 10 after a residual jump, Ng step only when
   max(gamma_s - beta v_s) < 40), convergence when max|gamma_new - gamma_s|
   < tol, and up to 4 attempts with halved mixing from h_gs0 on instability.
   Returns the final gamma_s (nsolve x nr) and, per solve, status 1
   (converged), 0 (maxiter reached) or -1 (unstable in every attempt: the
   caller continues in Gamma on the host), the iterations of the last
   attempt and the last residual.  The call keeps the DST-I table of the
   last nr (16 (nr+1) bytes) on the device for the life of the process,
   and the 8 nr^2-byte DST-I matrix as well when nsolve > 192 or
   nr > 8192.  Calls from several host threads run one after another.  No
   counterpart as one call in the reference (it solves one state point at a
   time on the host). */
int hy_hnc_iterate(int nsolve, int nr, double dr, double dk, const double* h_n0, const double* h_bvs,
                   const double* h_bvlk, const double* h_gs0, double tol, double mix, int maxiter, double* h_gs,
                   int* h_status, int* h_niter, double* h_resid);

#ifdef __cplusplus
}
#endif
#endif /* HYDFT_H */

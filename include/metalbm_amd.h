/*
 * metalbm_amd.h -- C ABI of the MI355X-native fused collide-and-stream-pull
 * step (libmetalbm_amd.so).
 *
 * The entry points replace, for one GPU and one process-local periodic
 * domain, the reference's per-step operator chain:
 *   Algorithm<...>::iterate            include/metaLBM/Algorithm.h:326-358
 *   Algorithm<Generic>::operator()     include/metaLBM/Algorithm.h:97-126
 *   Collision_ / ForcingScheme_ / Force_ / Equilibrium_ template plugins
 *                                      include/metaLBM/Collision.h:913-914,
 *                                      ForcingScheme.h:200, Force.h:628,
 *                                      Equilibrium.h:330
 *   initDistribution                   include/metaLBM/Initialize.h:91-117
 *   storeFields                        include/metaLBM/Algorithm.h:150-194
 *   TotalEnergy / reduce               include/metaLBM/Analysis.h:33-98,
 *                                      Communication.h:76-101
 * Plain C types only; every call returns LBM_OK (0) or a negative LBM_E*
 * code and records a message readable with lbm_last_error().  Nothing ever
 * calls exit() (the reference exits, Commons.h:6-38).
 *
 * Layout of host arrays: populations f[Q][N] (population-major SoA), vector
 * fields u[D][N], scalars rho[N]; cell = (x*ny + y)*nz + z in 3D (z fastest)
 * and x*ny + y in 2D (y fastest) -- the reference's halo-space order
 * (Domain.h:204-207, 272-276) without halos or FFTW padding.
 *
 * Threading: a handle belongs to one host thread.  All device work is
 * enqueued on the handle's HIP stream; calls that return data to the host
 * synchronise that stream, lbm_step does not.
 */
#ifndef METALBM_AMD_H
#define METALBM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

/* Error codes. */
enum {
  LBM_OK = 0,
  LBM_E_INVALID = -1,   /* bad argument or unsupported configuration */
  LBM_E_HIP = -2,       /* HIP runtime error (message has the HIP text) */
  LBM_E_NOMEM = -3,     /* device allocation failed */
  LBM_E_STATE = -4      /* call not valid in the handle's current state */
};

/* Enumerator values equal the reference's enum classes
 * (include/metaLBM/Options.h:12-41). */
enum { LBM_D2Q9 = 3, LBM_D3Q19 = 9, LBM_D3Q27 = 10 };             /* LatticeType */
enum { LBM_BGK = 1, LBM_ELBM = 2 };                                /* CollisionType */
enum { LBM_FS_NONE = 1, LBM_FS_GUO = 2, LBM_FS_SHANCHEN = 3,
       LBM_FS_EDM = 4 };                                           /* ForcingSchemeType */
enum { LBM_FORCE_NONE = 0, LBM_FORCE_CONSTANT = 4,
       LBM_FORCE_SINUSOIDAL = 5, LBM_FORCE_KOLMOGOROV = 6,
       LBM_FORCE_CONSTANT_SHELL = 7, LBM_FORCE_ENERGY_REMOVAL = 8,
       LBM_FORCE_TURBULENT_2D = 9 };                               /* ForceType */
enum { LBM_EQ_EXACT = 1, LBM_EQ_MA3 = 3 };                         /* EquilibriumType */

/* Arithmetic types (storage of populations, compute of the step). */
enum { LBM_F64 = 0, LBM_F32 = 1, LBM_F16 = 2 };

/* Kernel shape of the BGK step. */
enum {
  LBM_KERNEL_AUTO = 0,    /* the vector kernel for D3Q19 with FP32 storage and
                             for FP64 storage with FP32 compute, when the
                             contiguous extent (nz in 3D, ny in 2D) is
                             divisible by V; else one cell per lane (the
                             faster shape measured for each) */
  LBM_KERNEL_CELL = 1,    /* one cell per lane */
  LBM_KERNEL_VECTOR = 2   /* V consecutive cells per lane, aligned vector
                             accesses; needs nz (3D) / ny (2D) divisible by
                             a supported V (the widest is chosen) */
};

/* Entropic alpha solver (ELBM only). */
enum {
  LBM_ALPHA_NEWTON = 0,   /* reference Newton: warm start from the stored
                             alpha, |step| <= 1e-8, <= 50 iterations
                             (Collision.h:328-375, EntropicStep.h:111-140) */
  LBM_ALPHA_LEARNED = 1,  /* start from the MLP's alpha and make one Newton
                             correction: accept the start when |G| <=
                             residual_tol, or the corrected point when its
                             Taylor remainder bound is <= residual_tol --
                             and only where a bound from the same evaluation
                             shows the reference Newton from the stored alpha
                             reaching that root; else plain Newton at the
                             residual tolerance from the stored alpha decides
                             (NEWTON_RESIDUAL's solver, whose alpha = 2
                             fallbacks are the reference's), in a second
                             launch over the cells so deferred.  The alpha
                             state is kept as for NEWTON. */
  LBM_ALPHA_WARM_RESIDUAL = 2,  /* the certified correction from the stored
                             alpha: the method without the network (on D3Q27
                             up to two certified steps).  Its iterates are
                             plain Newton's, so it reaches the alpha
                             NEWTON_RESIDUAL reaches; the certificate saves
                             the last evaluation */
  LBM_ALPHA_NEWTON_RESIDUAL = 3  /* plain Newton at the residual tolerance: the
                             reference Newton (EntropicStep.h:111-140) from the
                             stored alpha, stopped when |G(alpha)| <=
                             residual_tol instead of |step| <= 1e-8, at most 50
                             evaluations, accepted inside (1, alphaMax), else
                             alpha = 2 (Collision.h:328-349).  Like the
                             reference's loop it bounds alpha only at
                             convergence, so it falls back exactly where the
                             reference Newton does; one launch */
};

typedef struct {
  int lattice;          /* LBM_D2Q9 | LBM_D3Q19 | LBM_D3Q27 */
  int collision;        /* LBM_BGK | LBM_ELBM */
  int forcing_scheme;   /* LBM_FS_* */
  int force;            /* LBM_FORCE_* */
  int equilibrium;      /* LBM_EQ_MA3 | LBM_EQ_EXACT (not for D3Q19) */
  int nx, ny, nz;       /* nz == 1 for D2Q9; nx*ny*nz < 2^31 */
  double tau;           /* relaxationTime (Input_prod.in:38) */
  double amplitude[3];  /* forceAmplitude (Input_prod.in:50) */
  double wavelength[3]; /* forceWaveLength (Input_prod.in:51) */
  int storage;          /* LBM_F64 | LBM_F32 | LBM_F16 (F16 needs shifted) */
  int shifted;          /* nonzero: store f_i - w_i instead of f_i */
  int compute;          /* LBM_F64 | LBM_F32 */
  int alpha_mode;       /* LBM_ALPHA_* (ELBM) */
  double residual_tol;  /* absolute H-residual tolerance |G(alpha)| of the
                           LEARNED, WARM_RESIDUAL and NEWTON_RESIDUAL solvers */
  int kernel;           /* LBM_KERNEL_* */
  /* Spectral forces (Force.h:294-623), a force field built by the library
   * (hipFFT, or a direct transform for a removal shell of <= 32 modes); shells are integer |k|^2 in [kmin^2, kmax^2]:
   *  CONSTANT_SHELL: F = curl psi, psi_hat = amplitude on the shell
   *    (MakeIncompressible, Transformer.h:301-389), set once;
   *  ENERGY_REMOVAL: F_d = -removal_amplitude_d (rho u_d) restricted to the
   *    removal shell, from the fields of the last storing step (before any:
   *    the initialiser's rho0 and u0, or after lbm_set_populations the
   *    restart fields rho = sum f_i, u = sum c_i f_i, undivided as in
   *    Initialize.h:124-145), recomputed before a step whenever those fields
   *    have changed (the result is the same as recomputing every step);
   *  TURBULENT_2D: the sum of both (injection: amplitude, force_kmin..kmax). */
  int force_kmin, force_kmax;     /* forcekMin, forcekMax (Input_prod_2d.in:51-52) */
  double removal_amplitude[3];    /* removalForceAmplitude (Input_prod_2d.in:54) */
  int removal_kmin, removal_kmax; /* removalForcekMin/kMax (Input_prod_2d.in:56-57) */
} lbm_config;

typedef struct lbm_handle lbm_handle;

/* Per-run entropic statistics (cells per branch, summed over steps). */
typedef struct {
  int64_t small;        /* max|fneq/f| <= 1e-3, alpha = 2      (Collision.h:356-359) */
  int64_t capped;       /* alphaMax < 2, alpha = 0.95 alphaMax (Collision.h:365-367) */
  int64_t newton;       /* root solved inside (1, alphaMax) */
  int64_t fallback;     /* Newton failed or out of bounds, alpha = 2 (Collision.h:344-346) */
  int64_t newton_iterations;  /* G evaluations (reference Newton iterations plus
                                 certified-correction evaluations) */
  int64_t learned_accepted;   /* LEARNED / WARM_RESIDUAL: solved cells accepted
                                 by the certified correction (no reference Newton) */
} lbm_elbm_stats;

int lbm_abi_version(void);

/* Fills a config with the reference defaults: D2Q9 BGK, Guo, no force,
 * Ma3, 16x16, tau 0.6, F64 storage and compute, unshifted; spectral-force
 * shells 1..2 and removal amplitude 1e-6 (Input_prod.in:52-58). */
int lbm_config_default(lbm_config *cfg);

/* Creates a handle on the current HIP device: validates cfg, allocates the
 * two population lattices (ping-pong) and, for ELBM, the alpha field set to 2
 * (Initialize.h:83-89).  Replaces Routine's constructor chain
 * (Routine.h:58-84) minus I/O. */
int lbm_create(const lbm_config *cfg, lbm_handle **out);
int lbm_destroy(lbm_handle *h);

/* Message of the last failed call on h (or of the last failed lbm_create
 * when h is NULL).  Valid until the next call on h. */
const char *lbm_last_error(const lbm_handle *h);

/* Sizes derived from the configuration. */
int lbm_sizes(const lbm_handle *h, int64_t *cells, int *q, int *d);
/* The handle's configuration (alpha_mode and residual_tol as last set;
 * kernel: the shape in use, LBM_KERNEL_CELL or LBM_KERNEL_VECTOR). */
int lbm_get_config(const lbm_handle *h, lbm_config *out);

/* The step kernel lbm_create chooses for cfg without creating anything:
 * *shape = LBM_KERNEL_CELL or LBM_KERNEL_VECTOR and *vec = the vector
 * kernel's cells per lane V (0 for the cell kernel).  Host-only (no device
 * is touched); LBM_E_INVALID for a configuration lbm_create rejects.  No
 * reference counterpart: the reference has one kernel shape
 * (Computation.cuh:39-51). */
int lbm_kernel_shape(const lbm_config *cfg, int *shape, int *vec);

/* f = feq(rho, u) (Initialize.h:106-117).  rho[N], u[D][N] on the host. */
int lbm_init_equilibrium(lbm_handle *h, const double *rho, const double *u);

/* f = feq(rho0, u0) with u0 seeded uniform noise in [-amp, amp) per
 * component: u_d(cell) = amp (2 U - 1), U = (SplitMix64(seed, cell*D + d)
 * >> 11) 2^-53.  Generated on the device. */
int lbm_init_noise(lbm_handle *h, double rho0, double amp, uint64_t seed);

/* f = feq(rho0, u0) with the Taylor-Green vortex u0 = U0 (sin x cos y cos z,
 * -cos x sin y cos z, 0) (2D: without the z factor), x = 2 pi i / n. */
int lbm_init_taylor_green(lbm_handle *h, double rho0, double u0);

/* f = feq(rho0, u0) with u0 a solenoidal random-phase field of shell
 * spectrum E(s) = s^4 exp(-2 (s/kpeak)^2), s = 1..kmax, and per-component rms
 * urms.  No reference counterpart: InitVelocityType::Decay is declared but
 * not implemented (Options.h:27).  Specification:
 *  - modes: integer k, |k_d| <= kmax (k_z = 0 in 2D); shell s(k) = nearest
 *    integer to |k|; active when 1 <= s <= kmax; N_s active modes in shell s;
 *  - the m-th active mode of the half space {k_x > 0} u {k_x = 0, k_y > 0} u
 *    {k_x = k_y = 0, k_z > 0}, in ascending (k_x, k_y, k_z) order, takes
 *    SplitMix64(seed) draws 3m, 3m+1, 3m+2 as angles theta1, theta2, phi =
 *    2 pi (r >> 11) 2^-53;
 *  - 3D: a = axis of the smallest |k_d| (ties: lowest d), e1 = k x a/|k x a|,
 *    e2 = k x e1/|k x e1|, uh = A_s (cos phi e^{i theta1} e1 +
 *    sin phi e^{i theta2} e2); 2D: uh = A_s e^{i theta1} (-k_y, k_x)/|k|;
 *    A_s = sqrt(E(s)/N_s); uh(-k) = conj(uh(k));
 *  - u(x) = gamma sum_k uh(k) e^{2 pi i (k_x x/nx + k_y y/ny + k_z z/nz)},
 *    gamma = urms sqrt(D / sum_s E(s)), so mean |u|^2 = D urms^2.
 * Requires 1 <= kmax and 2 kmax < every extent (LBM_E_INVALID otherwise). */
int lbm_init_spectral(lbm_handle *h, double rho0, double urms, int kmax,
                      double kpeak, uint64_t seed);

/* Populations as FP64 f_i (unshifted) on the host, [Q][N]. */
int lbm_set_populations(lbm_handle *h, const double *f);
int lbm_get_populations(lbm_handle *h, double *f);

/* The per-cell alpha field (ELBM warm start; Algorithm.h:103-106). */
int lbm_set_alpha(lbm_handle *h, const double *alpha);
int lbm_get_alpha(lbm_handle *h, double *alpha);

/* Advances nsteps fused collide-and-stream-pull steps (Routine.h:121-147,
 * Algorithm.h:326-358).  If store_last is nonzero the last step also writes
 * the output fields rho, u_hydro, alpha, force as storeFields does
 * (Algorithm.h:150-194).  Asynchronous on the handle's stream. */
int lbm_step(lbm_handle *h, int nsteps, int store_last);

/* Copies the fields written by the last storing step: rho[N],
 * u_hydro[D][N], alpha[N], force[D][N]; any pointer may be NULL. */
int lbm_get_fields(lbm_handle *h, double *rho, double *u_hydro,
                   double *alpha, double *force);

/* FP64 diagnostics of the stored fields, deterministic two-pass reduction:
 * mass = sum rho; energy = (1/N) sum 1/2 rho |u_h|^2 (Analysis.h:53-61);
 * enstrophy = (1/N) sum 1/2 |curl u_h|^2 with second-order central
 * differences and periodic wrap.  Any output pointer may be NULL. */
int lbm_reduce(lbm_handle *h, double *mass, double *energy, double *enstrophy);

/* lbm_reduce without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * the same results, bit for bit, with the partial sums in a static device
 * array of the library (1.5 MiB per GPU, held until the process ends) instead
 * of the buffer lbm_reduce allocates on first use.  Grids of more than 2^29
 * cells exceed that array and are rejected (LBM_E_INVALID).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (two host
 * threads, or handles on different streams) race. */
int lbm_reduce_noalloc(lbm_handle *h, double *mass, double *energy,
                       double *enstrophy);

/* ---- Bandwidth probes ------------------------------------------------------
 * Measurement aid, no reference counterpart: enqueues nsteps launches of a
 * probe that has the BGK vector kernel's memory access pattern -- 16 bytes
 * per lane, its tile order, per-row buffer resources, non-temporal loads
 * and stores --
 * and none of its arithmetic, over the handle's two lattices (ping-pong), so
 * that the kernel's achieved bandwidth can be read against what the same
 * pattern reaches on this GPU.  Kinds:
 *   LBM_PROBE_PULL_COPY  the Q reads at the pull source rows and the Q writes
 *                        of the step (2 Q streams);
 *   LBM_PROBE_READ       the Q reads only;
 *   LBM_PROBE_WRITE      the Q writes only;
 *   LBM_PROBE_COPY       one contiguous copy of the same 2 Q N sizeof(S)
 *                        bytes (one read and one write stream).
 * Needs the contiguous extent (nz in 3D, ny in 2D) to hold a multiple of 16
 * bytes.  Overwrites the populations: initialise again before stepping. */
enum { LBM_PROBE_PULL_COPY = 0, LBM_PROBE_READ = 1, LBM_PROBE_WRITE = 2,
       LBM_PROBE_COPY = 3 };
int lbm_bandwidth_probe(lbm_handle *h, int kind, int nsteps);

/* Entropic statistics accumulated since creation or the last reset. */
int lbm_get_elbm_stats(lbm_handle *h, lbm_elbm_stats *out);
int lbm_reset_elbm_stats(lbm_handle *h);

/* ---- Learned alpha (ELBM on D2Q9 and D3Q27) ------------------------------
 * No counterpart in the reference, which solves alpha by Newton only
 * (Collision.h:328-349).  A ReLU MLP (D2Q9: 9-32-32-1, D3Q27: 27-64-64-1,
 * FP32 master weights, trained by the library's GPU kernels) maps a cell's
 * normalised non-equilibrium moments to a first alpha.
 * Parameter layout: W1[h1][nin] b1[h1] W2[h2][h1] b2[h2] W3[h2] b3.
 * Features: m_k = (1/rho) sum_i fneq_i prod_d phi_{a_kd}(c_id) over the
 * tensor basis {1, c, c^2}^D, k = digits (a_x a_y [a_z]) in base 3. */
int lbm_mlp_shape(const lbm_handle *h, int *nin, int *h1, int *h2, int *nparams);
/* Allocates the network and initialises it (He-uniform, SplitMix64 seed). */
int lbm_mlp_init(lbm_handle *h, uint64_t seed);
int lbm_mlp_set_params(lbm_handle *h, const float *params);
int lbm_mlp_get_params(lbm_handle *h, float *params);
/* z_k = (m_k - in_mean_k) in_scale_k; alpha = out_mean + out_scale y. */
int lbm_mlp_set_normalization(lbm_handle *h, const float *in_mean,
                              const float *in_scale, double out_mean,
                              double out_scale);
int lbm_mlp_get_normalization(lbm_handle *h, float *in_mean, float *in_scale,
                              double *out_mean, double *out_scale);
/* Gradient of L = 1/(2n) sum (y - t)^2, t = (alpha - out_mean)/out_scale,
 * over the n host samples x[n][nin], computed by the training kernels. */
int lbm_mlp_gradient(lbm_handle *h, const float *x, const double *alpha, int n,
                     float *grad, double *loss);
/* alpha_hat for n host feature rows, computed by the step kernel's own
 * inference code (matrix cores: f16 operands for D2Q9, bf16 for D3Q27, FP32
 * accumulation, FP32 biases). */
int lbm_mlp_predict(lbm_handle *h, const float *x, int n, double *alpha);
/* Logs (features, alpha) of a `fraction` of the cells the Newton solver
 * solves during subsequent steps, up to `capacity` samples, into device
 * memory: every solve in LBM_ALPHA_NEWTON and LBM_ALPHA_NEWTON_RESIDUAL, in
 * the certified modes the solves of the cells their correction deferred
 * (data to retrain on).  Which cells are sampled is a fixed hash of the cell,
 * the salt and the step; when more solves qualify than `capacity` holds, the
 * ones kept depend on the order the wavefronts reach the log, so choose
 * `fraction` for the expected count to fit when runs must repeat exactly. */
int lbm_sample_log_enable(lbm_handle *h, int64_t capacity, double fraction,
                          uint32_t salt);
int lbm_sample_log_disable(lbm_handle *h);
int lbm_sample_log_count(lbm_handle *h, int64_t *n);
int lbm_sample_log_copy(lbm_handle *h, float *x, double *alpha, int64_t n);
/* Trains on the logged samples on the GPU, starting from the current
 * parameters (lbm_mlp_init(seed) first if none were initialised or loaded):
 * puts the samples in a canonical order (by alpha, then features; the log's
 * own order depends on wavefront timing), sets the normalisation from them,
 * then runs `iterations` Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on
 * minibatches of `batch` drawn by a seeded shuffle: the same samples and seed
 * give the same network.  Mean losses of the
 * first and last step are returned (either pointer may be NULL). */
int lbm_mlp_train(lbm_handle *h, int iterations, int batch, double lr,
                  uint64_t seed, double *loss_first, double *loss_last);
/* Switches the ELBM alpha solver (LBM_ALPHA_*) and its residual tolerance. */
int lbm_set_alpha_mode(lbm_handle *h, int mode, double residual_tol);

/* Stream control: the handle enqueues on `stream` (a hipStream_t, may be
 * NULL for the null stream) from now on. */
int lbm_set_stream(lbm_handle *h, void *stream);
int lbm_synchronize(lbm_handle *h);

/* Device pointer of the lattice holding the current populations and its
 * element size, for zero-copy callers. */
int lbm_device_populations(lbm_handle *h, void **ptr, int *elem_bytes);

/* Steps taken since creation. */
int64_t lbm_steps_done(const lbm_handle *h);

/* ---- HDF5 files in the reference layout ----------------------------------
 * One FP64 dataset per component, dimensions the global extents with the
 * last one padded to 2 (n/2 + 1) (gSD::pLength(), Domain.h:53-57,
 * MathVector.h:331-344); padding written as zero, ignored on reading.
 * libhdf5 (>= 1.10) is loaded at first use (dlopen: $METALBM_HDF5_LIBRARY,
 * libhdf5.so, libhdf5.so.103, /opt/conda/lib); LBM_E_STATE if absent. */
/* Fields of the last storing step, in FieldList::writeFields order
 * (FieldList.h:65-86, Routine.h:202-210, Writer.h:309-360): density,
 * velocity{X,Y[,Z]}, then as lbm_set_field_output selects: alpha (2 for BGK,
 * Collision.h:114-123), force{X,Y[,Z]}, and the vorticity of
 * Curl::executeSpace + normalize (Routine.h:129-132; lbm_vorticity with
 * LBM_SPECTRAL_REFERENCE): `vorticity` in 2D, vorticity{X,Y,Z} in 3D.
 * Writes the XDMF sidecar of FieldWriter<XDMF> (Writer.h:451-560) to
 * xdmf_path unless it is NULL. */
int lbm_write_fields_hdf5(lbm_handle *h, const char *path, const char *xdmf_path);
/* writeAlpha, writeForce, writeVorticity of the reference's input file
 * (src/Input_prod.in:68-72): which optional datasets lbm_write_fields_hdf5
 * writes.  All three are on for a new handle, as in every shipped input. */
int lbm_set_field_output(lbm_handle *h, int write_alpha, int write_force,
                         int write_vorticity);
/* Unshifted FP64 populations as datasets distribution0..distribution{Q-1}
 * (DistributionWriter, Writer.h:385-440, Routine.h:212-216). */
int lbm_write_distribution_hdf5(lbm_handle *h, const char *path);
/* Restart: reads distribution0..{Q-1} of the padded extents of this handle's
 * grid into the populations (Reader.h:106-158, Initialize.h:119-145). */
int lbm_read_distribution_hdf5(lbm_handle *h, const char *path);

/* ---- Spectral diagnostics of the stored fields ---------------------------
 * Forward/inverse transforms by hipFFT on the device; Fourier-space work by
 * the library's kernels.  Scalings:
 *  LBM_SPECTRAL_REFERENCE: the reference's numbers -- integer wavenumbers,
 *    vorticity divided by N twice and (2D) its imaginary part dropped
 *    (Transformer.h:100-111, 114-295), shells floor(|k|), weight 1/2 on the
 *    k_last = 0 plane, energy spectrum / N, forcing spectrum unnormalised
 *    (Analysis.h:100-176, AnalysisList.h:98-202);
 *  LBM_SPECTRAL_PHYSICAL: wavenumbers 2 pi k/n, the exact spectral curl,
 *    shells round(|k|), weight 1/2 on the k_last = 0 and n/2 planes, both
 *    spectra / N^2 (sum_s E(s) = mean 1/2 |u|^2 below the last shell).
 * Both zero the derivative along an even axis at its Nyquist index. */
enum { LBM_SPECTRAL_REFERENCE = 0, LBM_SPECTRAL_PHYSICAL = 1 };
/* Number of shells s = 0 .. max(n)/2 - 1 (gFD::maxWaveNumber,
 * FourierDomain.h:76-78). */
int lbm_spectral_shells(const lbm_handle *h, int *nshells);
/* Curl of u_hydro: omega[2D-3][N] (Curl::executeSpace + normalize,
 * Routine.h:129-132) and enstrophy = (1/N) sum 1/2 |omega|^2
 * (TotalEnstrophy, Analysis.h:68-98); either pointer may be NULL. */
int lbm_vorticity(lbm_handle *h, int scaling, double *omega, double *enstrophy);
/* Shell spectra of u_hydro (energy) and of the force field (forcing),
 * each [nshells] (PowerSpectra, SpectralAnalysisList::writeAnalyses);
 * either pointer may be NULL. */
int lbm_spectra(lbm_handle *h, int scaling, double *energy, double *forcing);

#ifdef __cplusplus
}
#endif

#endif /* METALBM_AMD_H */

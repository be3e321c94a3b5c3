/*
 * cpu_ref.h -- FP64 CPU restatement of metaLBM's fused collide-and-stream-pull
 * step.  TEST INFRASTRUCTURE ONLY: this oracle is the checker that the HIP
 * library is compared against.  Product code never links, loads or calls it.
 *
 * Every function restates the reference semantics of the files cited next to
 * it (paths relative to the metaLBM source tree), including the quirks that
 * parity has to reproduce:
 *   - Guo / Shan-Chen evaluate feq at the raw (unshifted) velocity
 *     (include/metaLBM/Collision.h:140-142, ForcingScheme.h:35-38);
 *   - Shan-Chen's collision source is zero (ForcingScheme.h:141-151);
 *   - the reported hydrodynamic velocity is u + F/(2 rho) except for None
 *     (ForcingScheme.h:26-33, 50-57);
 *   - ELBM warm-starts Newton from the previous step's alpha and returns the
 *     iterate *before* the final step (EntropicStep.h:111-140).
 *
 * Layout: unpadded SoA, population-major: f[iQ * N + cell] with
 * cell = (x * ny + y) * nz + z (3D, z fastest) or x * ny + y (2D, y fastest),
 * i.e. the reference's hSD order (Domain.h:204-207, 272-276) without halos or
 * FFTW padding.  Periodicity is modular source indexing, equivalent to the
 * reference's halo fill (Communication.h:134-180, Boundary.h:45-102).
 *
 * Parity status of each function is stated in its comment in cpu_ref.c.
 */
#ifndef METALBM_ORACLE_CPU_REF_H
#define METALBM_ORACLE_CPU_REF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Enumerator values are those of the reference enum classes
 * (include/metaLBM/Options.h:12-41), so one integer means the same thing on
 * both sides of a comparison. */
enum {
  REF_LATTICE_D2Q9 = 3,
  REF_LATTICE_D3Q19 = 9,
  REF_LATTICE_D3Q27 = 10
};
enum { REF_COLLISION_BGK = 1, REF_COLLISION_ELBM = 2 };
enum {
  REF_FS_NONE = 1,
  REF_FS_GUO = 2,
  REF_FS_SHANCHEN = 3,
  REF_FS_EDM = 4
};
enum {
  REF_FORCE_NONE = 0,
  REF_FORCE_CONSTANT = 4,
  REF_FORCE_SINUSOIDAL = 5,
  REF_FORCE_KOLMOGOROV = 6,
  REF_FORCE_CONSTANT_SHELL = 7,   /* field forces: supplied to ref_step_field */
  REF_FORCE_ENERGY_REMOVAL = 8,
  REF_FORCE_TURBULENT_2D = 9
};
enum { REF_EQ_EXACT = 1, REF_EQ_MA3 = 3 };

/* ELBM alpha branch taken for a cell (Collision.h:351-375). */
enum {
  REF_ALPHA_SMALL = 0,     /* max|fneq/f| <= 1e-3 -> alpha = 2        */
  REF_ALPHA_CAPPED = 1,    /* alphaMax < 2 -> alpha = 0.95 alphaMax    */
  REF_ALPHA_NEWTON = 2,    /* Newton converged inside (1, alphaMax)    */
  REF_ALPHA_FALLBACK = 3   /* Newton failed or left bounds -> alpha=2  */
};

typedef struct {
  int lattice, collision, forcing_scheme, force, equilibrium;
  int nx, ny, nz;          /* nz must be 1 for D2Q9 */
  double tau;
  double amplitude[3];
  double wavelength[3];
} ref_config;

typedef struct {
  int64_t branch[4];       /* cells per REF_ALPHA_* branch */
  int64_t newton_iterations;
} ref_stats;

/* Lattice table access (Lattice.h:145-210, 535-612, 614-703). */
int ref_lattice_q(int lattice);
int ref_lattice_d(int lattice);
int ref_lattice_c(int lattice, int iq, int d);
double ref_lattice_w(int lattice, int iq);

int64_t ref_cells(const ref_config *cfg);

/* Equilibrium_::calculate for one population (Equilibrium.h:17-33, 60-81,
 * 106-126).  u has D components. */
double ref_equilibrium(const ref_config *cfg, double rho, const double *u,
                       int iq);

/* Force_::setForce at local position (x, y, z) (Force.h:104-292). */
void ref_force_at(const ref_config *cfg, int x, int y, int z, double *force);

/* f = feq(rho, u) for every cell (Initialize.h:106-117).
 * rho[N], u[D*N] (component-major), f[Q*N]. */
void ref_init_equilibrium(const ref_config *cfg, const double *rho,
                          const double *u, double *f);

/* One fused step: next = collideAndStream(pull(prev)) for all cells
 * (Algorithm.h:97-126).  alpha[N] is the per-cell alpha field, read as the
 * Newton warm start and overwritten (Algorithm.h:103-106); it may be NULL
 * for BGK.  rho, u_hydro, force_out (any may be NULL) receive the fields the
 * reference stores on output steps (Algorithm.h:150-194).  stats may be
 * NULL.  Returns 0 or a negative error code. */
int ref_step(const ref_config *cfg, const double *prev, double *next,
             double *alpha, double *rho, double *u_hydro, double *force_out,
             ref_stats *stats);

/* ref_step with the force read per cell from force_field[D][N] (Generic
 * setForce, Force.h:39-48) instead of the analytic forces; required for the
 * spectral forces (ConstantShell, EnergyRemoval, Turbulent2D), whose field
 * the caller computes (oracle/spectral.py).  force_field may be NULL for the
 * analytic forces. */
int ref_step_field(const ref_config *cfg, const double *prev, double *next,
                   double *alpha, const double *force_field, double *rho,
                   double *u_hydro, double *force_out, ref_stats *stats);

/* Run nsteps with a two-lattice ping-pong; the result ends in f.
 * scratch must hold Q*N doubles.  Fields (may be NULL) are those of the last
 * step.  Returns 0 or a negative error code. */
int ref_run(const ref_config *cfg, double *f, double *scratch, double *alpha,
            int nsteps, double *rho, double *u_hydro, double *force_out,
            ref_stats *stats);

/* ELBM alpha for one cell from pulled populations f[Q] and fneq[Q]
 * (Collision.h:227-241, 284-375; EntropicStep.h:31-62, 111-140).
 * branch and iterations may be NULL. */
double ref_elbm_alpha(int lattice, const double *f, const double *fneq,
                      double alpha_guess, int *branch, int *iterations);

/* H-difference G(alpha) and its derivative (EntropicStep.h:31-62). */
double ref_entropic_g(int lattice, const double *f, const double *fneq,
                      double alpha);
double ref_entropic_dg(int lattice, const double *f, const double *fneq,
                       double alpha);

/* SplitMix64 (Steele, Lea, Flood 2014): the k-th output (k >= 0) of the
 * stream seeded with `seed`. */
uint64_t ref_splitmix64(uint64_t seed, uint64_t k);

/* Seeded uniform velocity noise in [-amp, amp) per component, cell-major
 * (x-major cells, component-minor draw order), written component-major:
 * u[d * N + cell] uses draw k = cell * D + d. */
void ref_init_noise_velocity(const ref_config *cfg, uint64_t seed, double amp,
                             double *u);

/* Taylor-Green vortex, node-centred x_i = 2 pi i / n:
 * u = U0 (sin x cos y cos z, -cos x sin y cos z, 0) in 3D,
 * u = U0 (sin x cos y, -cos x sin y) in 2D. */
void ref_init_taylor_green(const ref_config *cfg, double u0, double *u);

/* Solenoidal random-phase velocity field with shell spectrum
 * E(s) = s^4 exp(-2 (s / kpeak)^2) on shells s = 1..kmax (no reference
 * counterpart: the reference declares InitVelocityType::Decay without an
 * implementation, Options.h:27).  Specification:
 *  - modes: integer k with |k_d| <= kmax (k_z = 0 in 2D); shell s(k) = the
 *    integer nearest |k|, i.e. (2s-1)^2 <= 4|k|^2 < (2s+1)^2; a mode is
 *    active when 1 <= s <= kmax; N_s = number of active modes of shell s;
 *  - half space H: k_x > 0, or k_x = 0 and k_y > 0, or k_x = k_y = 0 and
 *    k_z > 0.  Active modes of H in ascending (k_x, k_y, k_z) order, m-th of
 *    them, take draws j = 3m, 3m+1, 3m+2 of ref_splitmix64(seed, j) as
 *    U = (r >> 11) 2^-53 and set theta1 = 2 pi U_0, theta2 = 2 pi U_1,
 *    phi = 2 pi U_2;
 *  - basis: 3D: a = axis of the smallest |k_d| (ties: lowest d),
 *    e1 = k x a / |k x a|, e2 = k x e1 / |k x e1|; 2D: e1 = (-k_y, k_x)/|k|;
 *  - amplitude: A_s = sqrt(E(s) / N_s); 3D: uh(k) = A_s (cos phi e^{i theta1} e1
 *    + sin phi e^{i theta2} e2); 2D: uh(k) = A_s e^{i theta1} e1;
 *    uh(-k) = conj(uh(k));
 *  - field: u(x) = gamma Re sum_k uh(k) exp(2 pi i (k_x x/nx + k_y y/ny + k_z z/nz)),
 *    gamma = urms sqrt(D / sum_s E(s)), so that mean |u|^2 = D urms^2 exactly
 *    (Parseval; requires 2 kmax < each extent).  Evaluated as three separable
 *    passes (sum over k_z, then k_y, then k_x, each in ascending order) with
 *    twiddles from tables of cos/sin(2 pi m / n).
 * Returns 0, or -1 when kmax < 1 or 2 kmax >= an extent. */
int ref_init_spectral_velocity(const ref_config *cfg, double urms, int kmax,
                               double kpeak, uint64_t seed, double *u);

/* Diagnostics on stored fields, FP64 with a fixed pairwise summation tree:
 *   mass = sum rho;  energy = (1/N) sum 1/2 rho |u|^2 (Analysis.h:53-61);
 *   enstrophy = (1/N) sum 1/2 |w|^2 with w the curl of u by second-order
 *   central differences and periodic wrap (this build's definition; the
 *   reference's spectral curl is not a physical enstrophy). */
void ref_reduce(const ref_config *cfg, const double *rho, const double *u,
                double *mass, double *energy, double *enstrophy);

/* Threads used by the OpenMP loops (1 if built without OpenMP). */
void ref_set_threads(int nthreads);
int ref_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif

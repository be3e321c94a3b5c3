/*
 * cpu_ref.c -- FP64 CPU restatement of metaLBM's fused collide-and-stream
 * step.  TEST INFRASTRUCTURE ONLY (see cpu_ref.h): the checker for the HIP
 * library, never part of the product path.
 *
 * Build with -ffp-contract=off so every multiply and add rounds exactly as
 * the reference's separate C++ operations do; the operation order of every
 * expression below follows the cited reference line.
 *
 * Parity pinning (what checks each function; tests/ holds the checks):
 *   lattice tables     isotropy sums to 4th order, opposite pairs, the
 *                      c_x=-1 / +1 block structure (tests/test_oracle_*.py)
 *   ref_equilibrium    moments of feq recover (rho, u); feq(u=0) = rho w;
 *                      Exact matches its closed product form
 *   ref_force_at       closed-form sin values at y = 0, lambda/4, lambda/2
 *   moments + stream   brute force against numpy roll of each population
 *   BGK collide        mass conservation, momentum gain (1-1/2tau) F per
 *                      step (Guo), rest state fixed point, tau = 1 identity
 *   ref_elbm_alpha     each branch forced with hand-built cells; Newton
 *                      root checked against a long-double bisection
 *   init / noise       SplitMix64 published test vectors
 *   ref_reduce         closed forms for uniform flow and Taylor-Green
 */
#include "cpu_ref.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Velocity sets, reference order (Lattice.h:188-206, 578-608, 657-699).     */

typedef struct {
  int q, d;
  int c[27][3];
  double w[27];
} lat_t;

static const lat_t LAT_D2Q9 = {
    9, 2,
    {{0, 0, 0}, {-1, 1, 0}, {-1, 0, 0}, {-1, -1, 0}, {1, -1, 0},
     {1, 0, 0}, {1, 1, 0}, {0, -1, 0}, {0, 1, 0}},
    {4.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 9.0,
     1.0 / 36.0, 1.0 / 9.0, 1.0 / 9.0}};

static const lat_t LAT_D3Q19 = {
    19, 3,
    {{0, 0, 0},  {-1, 0, 0}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
     {-1, 0, 1}, {1, 0, 0},  {1, 1, 0},   {1, -1, 0}, {1, 0, 1},
     {1, 0, -1}, {0, -1, 0}, {0, 0, -1},  {0, -1, -1}, {0, -1, 1},
     {0, 1, 0},  {0, 0, 1},  {0, 1, 1},   {0, 1, -1}},
    {1.0 / 3.0,  1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
     1.0 / 36.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
     1.0 / 36.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0,
     1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0}};

static const lat_t LAT_D3Q27 = {
    27, 3,
    {{0, 0, 0},   {-1, 0, 0},  {-1, -1, 0}, {-1, 1, 0},  {-1, 0, -1},
     {-1, 0, 1},  {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
     {1, 0, 0},   {1, 1, 0},   {1, -1, 0},  {1, 0, 1},   {1, 0, -1},
     {1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1}, {0, -1, 0},
     {0, 0, -1},  {0, -1, -1}, {0, -1, 1},  {0, 1, 0},   {0, 0, 1},
     {0, 1, 1},   {0, 1, -1}},
    {8.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
     1.0 / 54.0,  1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
     2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
     1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 2.0 / 27.0,
     2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  2.0 / 27.0,  2.0 / 27.0,
     1.0 / 54.0,  1.0 / 54.0}};

static const lat_t *lattice_of(int lattice) {
  switch (lattice) {
    case REF_LATTICE_D2Q9: return &LAT_D2Q9;
    case REF_LATTICE_D3Q19: return &LAT_D3Q19;
    case REF_LATTICE_D3Q27: return &LAT_D3Q27;
    default: return NULL;
  }
}

int ref_lattice_q(int lattice) {
  const lat_t *L = lattice_of(lattice);
  return L ? L->q : -1;
}
int ref_lattice_d(int lattice) {
  const lat_t *L = lattice_of(lattice);
  return L ? L->d : -1;
}
int ref_lattice_c(int lattice, int iq, int d) {
  const lat_t *L = lattice_of(lattice);
  return L ? L->c[iq][d] : 0;
}
double ref_lattice_w(int lattice, int iq) {
  const lat_t *L = lattice_of(lattice);
  return L ? L->w[iq] : 0.0;
}

int64_t ref_cells(const ref_config *cfg) {
  return (int64_t)cfg->nx * cfg->ny * cfg->nz;
}

static int ref_threads = 1;
void ref_set_threads(int n) { ref_threads = n > 0 ? n : 1; }
int ref_get_threads(void) { return ref_threads; }

static int check_config(const ref_config *cfg) {
  const lat_t *L = lattice_of(cfg->lattice);
  if (!L) return -1;
  if (cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0) return -2;
  if (L->d == 2 && cfg->nz != 1) return -2;
  if (cfg->collision != REF_COLLISION_BGK &&
      cfg->collision != REF_COLLISION_ELBM)
    return -3;
  if (cfg->forcing_scheme < REF_FS_NONE || cfg->forcing_scheme > REF_FS_EDM)
    return -4;
  if (cfg->force != REF_FORCE_NONE && cfg->force != REF_FORCE_CONSTANT &&
      cfg->force != REF_FORCE_SINUSOIDAL && cfg->force != REF_FORCE_KOLMOGOROV &&
      cfg->force != REF_FORCE_CONSTANT_SHELL && cfg->force != REF_FORCE_ENERGY_REMOVAL &&
      cfg->force != REF_FORCE_TURBULENT_2D)
    return -5;
  /* There is no Exact equilibrium for D3Q19 (Equilibrium.h:36-126). */
  if (cfg->equilibrium == REF_EQ_EXACT && cfg->lattice == REF_LATTICE_D3Q19)
    return -6;
  if (cfg->equilibrium != REF_EQ_EXACT && cfg->equilibrium != REF_EQ_MA3)
    return -6;
  if (!(cfg->tau > 0.0)) return -7;
  if (cfg->force == REF_FORCE_SINUSOIDAL)
    for (int d = 0; d < L->d; ++d)
      if (cfg->wavelength[d] == 0.0) return -8;
  if (cfg->force == REF_FORCE_KOLMOGOROV && cfg->wavelength[0] == 0.0)
    return -8;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Equilibrium (Equilibrium.h).                                              */

/* TruncationMa3 (Equilibrium.h:17-33); the terms are accumulated left to
 * right exactly as the reference expression is written, with
 * Power<T,n>::Do expanded to repeated multiplies (Helpers.h:94-114). */
static inline double feq_ma3(const lat_t *L, double rho, const double *u,
                             double u2, int i) {
  double cu = (double)L->c[i][0] * u[0];
  for (int d = 1; d < L->d; ++d) cu += (double)L->c[i][d] * u[d];
  double fe = 1.0 + cu * 3.0 - 0.5 * u2 * 3.0 + 0.5 * 9.0 * cu * cu -
              0.5 * 9.0 * cu * u2 + cu * cu * cu * 27.0 / 6.0 +
              0.125 * u2 * u2 * 9.0 - 0.25 * cu * cu * u2 * 27.0 +
              cu * cu * cu * cu * 81.0 / 24.0;
  return rho * L->w[i] * fe;
}

/* Exact product form for D2Q9 / D3Q27 (Equilibrium.h:60-81, 106-126) with
 * PowerBase's exponent in {-1, 0, 1} (Helpers.h:61-71). */
static inline double feq_exact(const lat_t *L, double rho, const double *u,
                               int i) {
  double fe = 1.0;
  for (int d = 0; d < L->d; ++d) {
    double s = sqrt(1.0 + 3.0 * u[d] * u[d]);
    double b = (2 * u[d] + s) / (1.0 - u[d]);
    int c = L->c[i][d];
    double pb = c == 1 ? b : (c == 0 ? 1.0 : 1.0 / b);
    fe *= (2.0 - s) * pb;
  }
  return rho * L->w[i] * fe;
}

static inline double feq_any(const lat_t *L, int eq, double rho,
                             const double *u, double u2, int i) {
  return eq == REF_EQ_EXACT ? feq_exact(L, rho, u, i)
                            : feq_ma3(L, rho, u, u2, i);
}

double ref_equilibrium(const ref_config *cfg, double rho, const double *u,
                       int iq) {
  const lat_t *L = lattice_of(cfg->lattice);
  double u2 = u[0] * u[0];
  for (int d = 1; d < L->d; ++d) u2 += u[d] * u[d];
  return feq_any(L, cfg->equilibrium, rho, u, u2, iq);
}

/* ------------------------------------------------------------------------ */
/* Forces (Force.h:104-292), evaluated at the local position without halo    */
/* or rank offset (Collision.h:86).                                          */

void ref_force_at(const ref_config *cfg, int x, int y, int z, double *F) {
  const lat_t *L = lattice_of(cfg->lattice);
  int pos[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) F[d] = 0.0;
  switch (cfg->force) {
    case REF_FORCE_CONSTANT:  /* Force.h:154-159 */
      for (int d = 0; d < L->d; ++d) F[d] = cfg->amplitude[d];
      break;
    case REF_FORCE_SINUSOIDAL:  /* Force.h:208-215: unsigned iP * 2, then * pi */
      for (int d = 0; d < L->d; ++d)
        F[d] = cfg->amplitude[d] *
               sin((double)((unsigned)pos[d] * 2u) * M_PI / cfg->wavelength[d]);
      break;
    case REF_FORCE_KOLMOGOROV:  /* Force.h:262-267: only F_x is ever set */
      F[0] = cfg->amplitude[0] *
             sin((double)((unsigned)pos[1] * 2u) * M_PI / cfg->wavelength[0]);
      break;
    default:  /* None: force = amplitude = 0 (Force.h:113-125) */
      break;
  }
}

/* ------------------------------------------------------------------------ */
/* Forcing schemes (ForcingScheme.h).                                        */

/* Guo source (ForcingScheme.h:99-117), raw velocity. */
static inline double guo_source(const lat_t *L, double guo_coef,
                                const double *u, const double *F, int i) {
  double cu = (double)L->c[i][0] * u[0];
  for (int d = 1; d < L->d; ++d) cu += (double)L->c[i][d] * u[d];
  double s3 = 3.0 * cu;
  double r = (((double)L->c[i][0] - u[0]) + (double)L->c[i][0] * s3) * F[0];
  for (int d = 1; d < L->d; ++d)
    r += (((double)L->c[i][d] - u[d]) + (double)L->c[i][d] * s3) * F[d];
  return guo_coef * L->w[i] * 3.0 * r;
}

/* Collision source for any scheme; feq_i is the equilibrium value the
 * collision passes in (ForcingScheme.h:68-78, 99-117, 141-151, 184-197). */
static inline double source(const lat_t *L, int scheme, int eq,
                            double guo_coef, double rho, const double *u,
                            const double *F, double feq_i, int i) {
  switch (scheme) {
    case REF_FS_GUO:
      return guo_source(L, guo_coef, u, F, i);
    case REF_FS_EDM: {
      double v[3] = {0.0, 0.0, 0.0};
      double k = 1.0 / rho;
      for (int d = 0; d < L->d; ++d) v[d] = u[d] + F[d] * k;
      double v2 = v[0] * v[0];
      for (int d = 1; d < L->d; ++d) v2 += v[d] * v[d];
      return feq_any(L, eq, rho, v, v2, i) - feq_i;
    }
    default:  /* None and Shan-Chen: zero source */
      return 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Entropic step (EntropicStep.h:31-62, 111-140; Collision.h:284-375).       */

static inline double entropic_g(const lat_t *L, const double *f,
                                const double *fneq, double alpha) {
  double g = 0.0;
  for (int i = 0; i < L->q; ++i) {
    double fm = f[i] - alpha * fneq[i];
    g += f[i] * log(f[i] / L->w[i]) - fm * log(fm / L->w[i]);
  }
  return g;
}

static inline double entropic_dg(const lat_t *L, const double *f,
                                 const double *fneq, double alpha) {
  double dg = 0.0;
  for (int i = 0; i < L->q; ++i) {
    double fm = f[i] - alpha * fneq[i];
    dg += fneq[i] * (1 + log(fm / L->w[i]));
  }
  return dg;
}

double ref_entropic_g(int lattice, const double *f, const double *fneq,
                      double alpha) {
  return entropic_g(lattice_of(lattice), f, fneq, alpha);
}
double ref_entropic_dg(int lattice, const double *f, const double *fneq,
                       double alpha) {
  return entropic_dg(lattice_of(lattice), f, fneq, alpha);
}

static inline double elbm_alpha(const lat_t *L, const double *f,
                                const double *fneq, double alpha_guess,
                                int *branch, int *iterations) {
  int its = 0;
  /* isDeviationSmall (Collision.h:284-303) */
  int small = 1;
  for (int i = 0; i < L->q; ++i) {
    double dev = fabs(fneq[i]) / f[i];
    if (dev > 1.0e-3) small = 0;
  }
  if (small) {
    *branch = REF_ALPHA_SMALL;
    *iterations = 0;
    return 2.0;
  }
  /* calculateAlphaMax (Collision.h:305-326) */
  double amax = 2.5;
  for (int i = 0; i < L->q; ++i) {
    if (fneq[i] > 0) {
      double t = fabs(f[i]) / fneq[i];
      if (t < amax) amax = t;
    }
  }
  if (amax < 2.) {
    *branch = REF_ALPHA_CAPPED;
    *iterations = 0;
    return 0.95 * amax;
  }
  /* solveAlpha -> NewtonRaphsonSolver (Collision.h:328-349,
   * EntropicStep.h:111-140): the returned alpha is the iterate at which
   * |step| <= 1e-8 was observed, before that step is applied. */
  const double tol = 1e-8;
  const double amin = 1.;
  double x = alpha_guess;
  double step = 0.0;
  int ok = 0;
  for (int it = 1; it <= 50; ++it) {
    x = x - step;
    its = it;
    double g = entropic_g(L, f, fneq, x);
    double dg = entropic_dg(L, f, fneq, x);
    step = g / dg;
    if (fabs(step) <= tol) {
      ok = (x > amin && x < amax);
      break;
    }
  }
  *iterations = its;
  if (!ok) {
    *branch = REF_ALPHA_FALLBACK;
    return 2.0;
  }
  *branch = REF_ALPHA_NEWTON;
  return x;
}

double ref_elbm_alpha(int lattice, const double *f, const double *fneq,
                      double alpha_guess, int *branch, int *iterations) {
  int b, its;
  double a = elbm_alpha(lattice_of(lattice), f, fneq, alpha_guess, &b, &its);
  if (branch) *branch = b;
  if (iterations) *iterations = its;
  return a;
}

/* ------------------------------------------------------------------------ */
/* The fused per-cell step (Algorithm.h:97-126).                             */

typedef struct {
  double omega_f;   /* 1 - 2 beta            (Collision.h:145) */
  double omega_eq;  /* 2 beta                (Collision.h:147) */
  double beta;      /* 1 / (2 tau)           (Collision.h:122) */
  double guo_coef;  /* 1 - 1 / (2 tau0)      (ForcingScheme.h:115) */
  const double *ffield;  /* force field [D][N] (Generic setForce) or NULL */
} coefs_t;

static inline int wrap(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

static inline void cell_step(const lat_t *L, const ref_config *cfg,
                             const coefs_t *k, const double *prev,
                             double *next, double *alpha, double *rho_out,
                             double *uh_out, double *F_out, int64_t N, int x,
                             int y, int z, int64_t *branch_count,
                             int64_t *newton_its) {
  const int Q = L->q, D = L->d;
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  const int64_t cell = ((int64_t)x * ny + y) * nz + z;
  double f[27];
  /* Pull: f_i = prev[x - c_i, i] with periodic wrap (Collision.h:146). */
  for (int i = 0; i < Q; ++i) {
    int xs = wrap(x - L->c[i][0], nx);
    int ys = wrap(y - L->c[i][1], ny);
    int zs = D == 3 ? wrap(z - L->c[i][2], nz) : 0;
    f[i] = prev[(int64_t)i * N + ((int64_t)xs * ny + ys) * nz + zs];
  }
  /* calculateMoments (Collision.h:72-79 -> Moment.h:13-47). */
  double rho = f[0];
  for (int i = 1; i < Q; ++i) rho += f[i];
  double u[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < D; ++d) u[d] = (double)L->c[0][d] * f[0];
  for (int i = 1; i < Q; ++i)
    for (int d = 0; d < D; ++d) u[d] += (double)L->c[i][d] * f[i];
  for (int d = 0; d < D; ++d) u[d] /= rho;
  double u2 = u[0] * u[0];
  for (int d = 1; d < D; ++d) u2 += u[d] * u[d];
  /* setForce (Collision.h:81-88). */
  double F[3];
  if (k->ffield) {
    /* Force<Generic>::setForce reads the force field (Force.h:39-48). */
    F[0] = F[1] = F[2] = 0.0;
    for (int d = 0; d < D; ++d) F[d] = k->ffield[(int64_t)d * N + cell];
  } else {
    ref_force_at(cfg, x, y, z, F);
  }
  const int eq = cfg->equilibrium, scheme = cfg->forcing_scheme;
  double a = 2.0;
  if (cfg->collision == REF_COLLISION_BGK) {
    /* BGK collideAndStream (Collision.h:134-151). */
    for (int i = 0; i < Q; ++i) {
      double fe = feq_any(L, eq, rho, u, u2, i);
      next[(int64_t)i * N + cell] =
          k->omega_f * f[i] + k->omega_eq * fe +
          source(L, scheme, eq, k->guo_coef, rho, u, F, fe, i);
    }
  } else {
    /* ELBM calculateRelaxationTime (Collision.h:227-241): fneq is parked in
     * the destination slots, alpha warm-starts from the stored field. */
    double fneq[27];
    for (int i = 0; i < Q; ++i) fneq[i] = f[i] - feq_any(L, eq, rho, u, u2, i);
    int b, its;
    a = elbm_alpha(L, f, fneq, alpha[cell], &b, &its);
    branch_count[b] += 1;
    *newton_its += its;
    double tau = 1.0 / (a * k->beta);
    double rtau = 1.0 / tau;
    /* ELBM collideAndStream (Collision.h:243-258). */
    for (int i = 0; i < Q; ++i) {
      double fe = f[i] - fneq[i];
      next[(int64_t)i * N + cell] =
          f[i] - rtau * fneq[i] +
          source(L, scheme, eq, k->guo_coef, rho, u, F, fe, i);
    }
  }
  if (alpha) alpha[cell] = a;  /* Algorithm.h:106 */
  /* storeFields (Algorithm.h:150-194). */
  if (rho_out) rho_out[cell] = rho;
  if (uh_out) {
    if (scheme == REF_FS_NONE) {
      for (int d = 0; d < D; ++d) uh_out[(int64_t)d * N + cell] = u[d];
    } else {
      double h = 0.5 / rho;
      for (int d = 0; d < D; ++d)
        uh_out[(int64_t)d * N + cell] = u[d] + F[d] * h;
    }
  }
  if (F_out)
    for (int d = 0; d < D; ++d) F_out[(int64_t)d * N + cell] = F[d];
}

static void step_lattice(const lat_t *L, const ref_config *cfg,
                         const double *prev, double *next, double *alpha,
                         const double *ffield, double *rho, double *uh,
                         double *F_out, ref_stats *stats) {
  coefs_t k;
  k.ffield = ffield;
  k.beta = 1.0 / (2.0 * cfg->tau);
  k.omega_f = 1. - 2. * k.beta;
  k.omega_eq = 2. * k.beta;
  k.guo_coef = 1.0 - 1.0 / (2.0 * cfg->tau);
  const int64_t N = ref_cells(cfg);
  int64_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, its = 0;
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
#pragma omp parallel for schedule(static) num_threads(ref_threads) \
    reduction(+ : b0, b1, b2, b3, its)
  for (int x = 0; x < nx; ++x) {
    int64_t bc[4] = {0, 0, 0, 0};
    int64_t it = 0;
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z)
        cell_step(L, cfg, &k, prev, next, alpha, rho, uh, F_out, N, x, y, z,
                  bc, &it);
    b0 += bc[0];
    b1 += bc[1];
    b2 += bc[2];
    b3 += bc[3];
    its += it;
  }
  if (stats) {
    stats->branch[0] += b0;
    stats->branch[1] += b1;
    stats->branch[2] += b2;
    stats->branch[3] += b3;
    stats->newton_iterations += its;
  }
}

int ref_step(const ref_config *cfg, const double *prev, double *next,
             double *alpha, double *rho, double *u_hydro, double *force_out,
             ref_stats *stats) {
  return ref_step_field(cfg, prev, next, alpha, NULL, rho, u_hydro, force_out,
                        stats);
}

int ref_step_field(const ref_config *cfg, const double *prev, double *next,
                   double *alpha, const double *force_field, double *rho,
                   double *u_hydro, double *force_out, ref_stats *stats) {
  int e = check_config(cfg);
  if (e) return e;
  if (cfg->collision == REF_COLLISION_ELBM && !alpha) return -9;
  /* The spectral forces live in a force field (Force.h:294-623). */
  if (cfg->force >= REF_FORCE_CONSTANT_SHELL && !force_field) return -10;
  /* Dispatch on a constant table so the compiler specialises the loops. */
  switch (cfg->lattice) {
    case REF_LATTICE_D2Q9:
      step_lattice(&LAT_D2Q9, cfg, prev, next, alpha, force_field, rho, u_hydro,
                   force_out, stats);
      break;
    case REF_LATTICE_D3Q19:
      step_lattice(&LAT_D3Q19, cfg, prev, next, alpha, force_field, rho, u_hydro,
                   force_out, stats);
      break;
    default:
      step_lattice(&LAT_D3Q27, cfg, prev, next, alpha, force_field, rho, u_hydro,
                   force_out, stats);
      break;
  }
  return 0;
}

int ref_run(const ref_config *cfg, double *f, double *scratch, double *alpha,
            int nsteps, double *rho, double *u_hydro, double *force_out,
            ref_stats *stats) {
  int e = check_config(cfg);
  if (e) return e;
  const int64_t n = ref_cells(cfg) * ref_lattice_q(cfg->lattice);
  double *a = f, *b = scratch;
  for (int s = 0; s < nsteps; ++s) {
    const int last = (s == nsteps - 1);
    e = ref_step(cfg, a, b, alpha, last ? rho : NULL, last ? u_hydro : NULL,
                 last ? force_out : NULL, stats);
    if (e) return e;
    double *t = a;
    a = b;
    b = t;
  }
  if (a != f) memcpy(f, a, (size_t)n * sizeof(double));
  return 0;
}

void ref_init_equilibrium(const ref_config *cfg, const double *rho,
                          const double *u, double *f) {
  const lat_t *L = lattice_of(cfg->lattice);
  const int64_t N = ref_cells(cfg);
#pragma omp parallel for schedule(static) num_threads(ref_threads)
  for (int64_t c = 0; c < N; ++c) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < L->d; ++d) v[d] = u[(int64_t)d * N + c];
    double v2 = v[0] * v[0];
    for (int d = 1; d < L->d; ++d) v2 += v[d] * v[d];
    for (int i = 0; i < L->q; ++i)
      f[(int64_t)i * N + c] = feq_any(L, cfg->equilibrium, rho[c], v, v2, i);
  }
}

/* ------------------------------------------------------------------------ */
/* Initial conditions (no reference counterpart: Initialize.h only has      */
/* homogeneous fields; these are the build's own definitions).              */

uint64_t ref_splitmix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void ref_init_noise_velocity(const ref_config *cfg, uint64_t seed, double amp,
                             double *u) {
  const int D = ref_lattice_d(cfg->lattice);
  const int64_t N = ref_cells(cfg);
  for (int64_t c = 0; c < N; ++c)
    for (int d = 0; d < D; ++d) {
      uint64_t r = ref_splitmix64(seed, (uint64_t)(c * D + d));
      double unit = (double)(r >> 11) * 0x1.0p-53; /* [0, 1) */
      u[(int64_t)d * N + c] = amp * (2.0 * unit - 1.0);
    }
}

/* sin and cos are evaluated in separate table loops: a compiler that sees
 * sin(X) and cos(X) together may merge them into sincos(), which glibc does
 * not round identically to sin() (e.g. X = 2 pi 19 / 20). */
static void trig_table(int n, double *s, double *c) {
  const double two_pi = 2.0 * M_PI;
  for (int i = 0; i < n; ++i) s[i] = sin(two_pi * i / n);
  for (int i = 0; i < n; ++i) c[i] = cos(two_pi * i / n);
}

void ref_init_taylor_green(const ref_config *cfg, double u0, double *u) {
  const int D = ref_lattice_d(cfg->lattice);
  const int64_t N = ref_cells(cfg);
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  double *t = (double *)malloc(sizeof(double) * 2 * (size_t)(nx + ny + nz));
  double *sx = t, *cx = sx + nx, *sy = cx + nx, *cy = sy + ny, *sz = cy + ny,
         *cz = sz + nz;
  trig_table(nx, sx, cx);
  trig_table(ny, sy, cy);
  trig_table(nz, sz, cz);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        int64_t c = ((int64_t)x * ny + y) * nz + z;
        double fz = D == 3 ? cz[z] : 1.0;
        u[c] = u0 * sx[x] * cy[y] * fz;
        u[N + c] = -u0 * cx[x] * sy[y] * fz;
        if (D == 3) u[2 * N + c] = 0.0;
      }
  free(t);
}

static int shell_of(int k2) {
  int s = (int)sqrt((double)k2);
  while ((2 * s + 1) * (2 * s + 1) <= 4 * k2) ++s;
  while (s > 0 && (2 * s - 1) * (2 * s - 1) > 4 * k2) --s;
  return s;
}

static void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

int ref_init_spectral_velocity(const ref_config *cfg, double urms, int kmax,
                               double kpeak, uint64_t seed, double *u) {
  const int D = ref_lattice_d(cfg->lattice);
  const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
  if (kmax < 1) return -1;
  for (int d = 0; d < D; ++d)
    if (2 * kmax >= n[d]) return -1;
  const int K = kmax, M = 2 * K + 1, Kz = D == 3 ? K : 0, Mz = 2 * Kz + 1;
  const size_t nm = (size_t)M * M * Mz;
  /* shell counts and amplitudes */
  int *cnt = (int *)calloc((size_t)K + 1, sizeof(int));
  for (int ix = 0; ix < M; ++ix)
    for (int iy = 0; iy < M; ++iy)
      for (int iz = 0; iz < Mz; ++iz) {
        const int kx = ix - K, ky = iy - K, kz = iz - Kz;
        const int s = shell_of(kx * kx + ky * ky + kz * kz);
        if (s >= 1 && s <= K) cnt[s]++;
      }
  double *amp = (double *)calloc((size_t)K + 1, sizeof(double));
  double esum = 0.0;
  for (int s = 1; s <= K; ++s) {
    const double r = s / kpeak;
    const double e = (double)s * s * s * s * exp(-2.0 * r * r);
    esum += e;
    amp[s] = cnt[s] > 0 ? sqrt(e / cnt[s]) : 0.0;
  }
  /* draws: angles of the active half-space modes, in order */
  size_t nh = 0;
  for (int ix = 0; ix < M; ++ix)
    for (int iy = 0; iy < M; ++iy)
      for (int iz = 0; iz < Mz; ++iz) {
        const int kx = ix - K, ky = iy - K, kz = iz - Kz;
        const int s = shell_of(kx * kx + ky * ky + kz * kz);
        const int half = kx > 0 || (kx == 0 && (ky > 0 || (ky == 0 && kz > 0)));
        if (s >= 1 && s <= K && half) nh++;
      }
  double *ang = (double *)malloc(sizeof(double) * 3 * (nh ? nh : 1));
  double *cs = (double *)malloc(sizeof(double) * 3 * (nh ? nh : 1));
  double *sn = (double *)malloc(sizeof(double) * 3 * (nh ? nh : 1));
  for (size_t j = 0; j < 3 * nh; ++j) {
    const double unit = (double)(ref_splitmix64(seed, (uint64_t)j) >> 11) * 0x1.0p-53;
    ang[j] = 2.0 * M_PI * unit;
  }
  /* cos and sin in separate loops (see trig_table) */
  for (size_t j = 0; j < 3 * nh; ++j) cs[j] = cos(ang[j]);
  for (size_t j = 0; j < 3 * nh; ++j) sn[j] = sin(ang[j]);
  /* mode cube uh[mode][d] (re, im) */
  double *ur = (double *)calloc(nm * D, sizeof(double));
  double *ui = (double *)calloc(nm * D, sizeof(double));
  size_t m = 0;
  for (int ix = 0; ix < M; ++ix)
    for (int iy = 0; iy < M; ++iy)
      for (int iz = 0; iz < Mz; ++iz) {
        const int kx = ix - K, ky = iy - K, kz = iz - Kz;
        const int s = shell_of(kx * kx + ky * ky + kz * kz);
        const int half = kx > 0 || (kx == 0 && (ky > 0 || (ky == 0 && kz > 0)));
        if (!(s >= 1 && s <= K && half)) continue;
        const double c1 = cs[3 * m], s1 = sn[3 * m];
        const double c2 = cs[3 * m + 1], s2 = sn[3 * m + 1];
        const double cp = cs[3 * m + 2], sp = sn[3 * m + 2];
        ++m;
        double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
        if (D == 3) {
          const double k[3] = {(double)kx, (double)ky, (double)kz};
          int a = 0;
          if (fabs(k[1]) < fabs(k[a])) a = 1;
          if (fabs(k[2]) < fabs(k[a])) a = 2;
          double ea[3] = {0, 0, 0}, e1[3], e2[3], w[3];
          ea[a] = 1.0;
          cross3(k, ea, w);
          double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          for (int d = 0; d < 3; ++d) e1[d] = w[d] / nw;
          cross3(k, e1, w);
          nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          for (int d = 0; d < 3; ++d) e2[d] = w[d] / nw;
          for (int d = 0; d < 3; ++d) {
            re[d] = amp[s] * (cp * c1 * e1[d] + sp * c2 * e2[d]);
            im[d] = amp[s] * (cp * s1 * e1[d] + sp * s2 * e2[d]);
          }
        } else {
          const double nk = sqrt((double)(kx * kx + ky * ky));
          const double e1[2] = {-ky / nk, kx / nk};
          for (int d = 0; d < 2; ++d) {
            re[d] = amp[s] * (c1 * e1[d]);
            im[d] = amp[s] * (s1 * e1[d]);
          }
        }
        const size_t p = ((size_t)ix * M + iy) * Mz + iz;
        const size_t q = ((size_t)(M - 1 - ix) * M + (M - 1 - iy)) * Mz + (Mz - 1 - iz);
        for (int d = 0; d < D; ++d) {
          ur[p * D + d] = re[d];
          ui[p * D + d] = im[d];
          ur[q * D + d] = re[d];
          ui[q * D + d] = -im[d];
        }
      }
  const double gamma = urms * sqrt((double)D / esum);
  /* twiddle tables */
  double *tc[3], *ts[3];
  for (int d = 0; d < 3; ++d) {
    tc[d] = (double *)malloc(sizeof(double) * n[d]);
    ts[d] = (double *)malloc(sizeof(double) * n[d]);
    trig_table(n[d], ts[d], tc[d]);
  }
  const int nx = n[0], ny = n[1], nz = n[2];
  /* pass z: t1[ix][iy][z][d] */
  double *t1r = (double *)malloc(sizeof(double) * (size_t)M * M * nz * D);
  double *t1i = (double *)malloc(sizeof(double) * (size_t)M * M * nz * D);
#pragma omp parallel for schedule(static) num_threads(ref_threads)
  for (int ixy = 0; ixy < M * M; ++ixy)
    for (int z = 0; z < nz; ++z)
      for (int d = 0; d < D; ++d) {
        double re = 0.0, im = 0.0;
        for (int iz = 0; iz < Mz; ++iz) {
          const int kz = iz - Kz;
          const int t = (int)((((int64_t)kz * z) % nz + nz) % nz);
          const double c = tc[2][t], sv = ts[2][t];
          const size_t p = ((size_t)ixy * Mz + iz) * D + d;
          const double a = ur[p], b = ui[p];
          const double tr = a * c - b * sv;
          const double ti = a * sv + b * c;
          re += tr;
          im += ti;
        }
        const size_t o = ((size_t)ixy * nz + z) * D + d;
        t1r[o] = re;
        t1i[o] = im;
      }
  /* pass y: t2[ix][y][z][d] */
  double *t2r = (double *)malloc(sizeof(double) * (size_t)M * ny * nz * D);
  double *t2i = (double *)malloc(sizeof(double) * (size_t)M * ny * nz * D);
#pragma omp parallel for schedule(static) num_threads(ref_threads)
  for (int ix = 0; ix < M; ++ix)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z)
        for (int d = 0; d < D; ++d) {
          double re = 0.0, im = 0.0;
          for (int iy = 0; iy < M; ++iy) {
            const int ky = iy - K;
            const int t = (int)((((int64_t)ky * y) % ny + ny) % ny);
            const double c = tc[1][t], sv = ts[1][t];
            const size_t p = (((size_t)ix * M + iy) * nz + z) * D + d;
            const double a = t1r[p], b = t1i[p];
            const double tr = a * c - b * sv;
            const double ti = a * sv + b * c;
            re += tr;
            im += ti;
          }
          const size_t o = (((size_t)ix * ny + y) * nz + z) * D + d;
          t2r[o] = re;
          t2i[o] = im;
        }
  /* pass x: u[d][cell] = gamma Re(...) */
  const int64_t N = ref_cells(cfg);
#pragma omp parallel for schedule(static) num_threads(ref_threads)
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z)
        for (int d = 0; d < D; ++d) {
          double re = 0.0;
          for (int ix = 0; ix < M; ++ix) {
            const int kx = ix - K;
            const int t = (int)((((int64_t)kx * x) % nx + nx) % nx);
            const double c = tc[0][t], sv = ts[0][t];
            const size_t p = (((size_t)ix * ny + y) * nz + z) * D + d;
            const double tr = t2r[p] * c - t2i[p] * sv;
            re += tr;
          }
          u[(int64_t)d * N + ((int64_t)x * ny + y) * nz + z] = gamma * re;
        }
  for (int d = 0; d < 3; ++d) {
    free(tc[d]);
    free(ts[d]);
  }
  free(t1r); free(t1i); free(t2r); free(t2i);
  free(ur); free(ui); free(ang); free(cs); free(sn); free(amp); free(cnt);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics.                                                              */

/* Pairwise sum with a fixed tree: deterministic and O(log n) error growth. */
static double pairwise_sum(const double *v, int64_t n) {
  if (n <= 8) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += v[i];
    return s;
  }
  int64_t h = n / 2;
  return pairwise_sum(v, h) + pairwise_sum(v + h, n - h);
}

void ref_reduce(const ref_config *cfg, const double *rho, const double *u,
                double *mass, double *energy, double *enstrophy) {
  const int D = ref_lattice_d(cfg->lattice);
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  const int64_t N = ref_cells(cfg);
  double *e = (double *)malloc((size_t)N * sizeof(double));
  double *w = (double *)malloc((size_t)N * sizeof(double));
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        int64_t c = ((int64_t)x * ny + y) * nz + z;
        double s = 0.0;
        for (int d = 0; d < D; ++d) s += 0.5 * rho[c] * u[d * N + c] * u[d * N + c];
        e[c] = s;
        /* Central differences with periodic wrap. */
#define U(d, X, Y, Z) u[(int64_t)(d) * N + ((int64_t)(X) * ny + (Y)) * nz + (Z)]
        int xp = wrap(x + 1, nx), xm = wrap(x - 1, nx);
        int yp = wrap(y + 1, ny), ym = wrap(y - 1, ny);
        if (D == 2) {
          double om = 0.5 * (U(1, xp, y, z) - U(1, xm, y, z)) -
                      0.5 * (U(0, x, yp, z) - U(0, x, ym, z));
          w[c] = 0.5 * om * om;
        } else {
          int zp = wrap(z + 1, nz), zm = wrap(z - 1, nz);
          double ox = 0.5 * (U(2, x, yp, z) - U(2, x, ym, z)) -
                      0.5 * (U(1, x, y, zp) - U(1, x, y, zm));
          double oy = 0.5 * (U(0, x, y, zp) - U(0, x, y, zm)) -
                      0.5 * (U(2, xp, y, z) - U(2, xm, y, z));
          double oz = 0.5 * (U(1, xp, y, z) - U(1, xm, y, z)) -
                      0.5 * (U(0, x, yp, z) - U(0, x, ym, z));
          w[c] = 0.5 * (ox * ox + oy * oy + oz * oz);
        }
#undef U
      }
  if (mass) *mass = pairwise_sum(rho, N);
  if (energy) *energy = pairwise_sum(e, N) / (double)N;
  if (enstrophy) *enstrophy = pairwise_sum(w, N) / (double)N;
  free(e);
  free(w);
}

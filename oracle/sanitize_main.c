/*
 * sanitize_main.c -- drives every entry point of the CPU oracle (cpu_ref.c,
 * mlp_ref.c) for an AddressSanitizer + UndefinedBehaviorSanitizer build
 * (`make -C oracle sanitize`, run by tests/test_sanitizers.py).  TEST
 * INFRASTRUCTURE ONLY.  Covers every lattice, collision, forcing scheme,
 * analytic force and equilibrium on small and degenerate extents (axes of
 * length 1 and 2), entropic cells in every alpha branch, the initialisers,
 * the reduction and the MLP reference.  Exits 0 when every call returned
 * normally; a sanitizer finding aborts the process.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "cpu_ref.h"

int mlp_ref_param_count(int nin, int h1, int h2);
void mlp_ref_forward(const double *p, int nin, int h1, int h2, int batch,
                     const double *x, double *y);
double mlp_ref_loss(const double *p, int nin, int h1, int h2, int batch,
                    const double *x, const double *t);
double mlp_ref_grad(const double *p, int nin, int h1, int h2, int batch,
                    const double *x, const double *t, double *g);
void mlp_ref_adam(double *p, const double *g, double *m, double *v, int n,
                  double lr, double b1, double b2, double eps, int t);

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

static int run_case(int lat, int col, int fs, int force, int eq, int nx, int ny, int nz,
                    double amp, int steps) {
  ref_config c = {lat, col, fs, force, eq, nx, ny, nz, col == REF_COLLISION_ELBM ? 0.5003 : 0.6,
                  {1e-4, -2e-4, 1e-4}, {4.0, 3.0, 5.0}};
  const int64_t N = ref_cells(&c);
  const int Q = ref_lattice_q(lat), D = ref_lattice_d(lat);
  double *rho = xcalloc((size_t)N, sizeof(double));
  double *u = xcalloc((size_t)N * D, sizeof(double));
  double *f = xcalloc((size_t)N * Q, sizeof(double));
  double *scratch = xcalloc((size_t)N * Q, sizeof(double));
  double *alpha = xcalloc((size_t)N, sizeof(double));
  double *uh = xcalloc((size_t)N * D, sizeof(double));
  double *F = xcalloc((size_t)N * D, sizeof(double));
  for (int64_t k = 0; k < N; ++k) {
    rho[k] = 1.0;
    alpha[k] = 2.0;
  }
  ref_init_noise_velocity(&c, 42, amp, u);
  ref_init_equilibrium(&c, rho, u, f);
  ref_stats st = {{0, 0, 0, 0}, 0};
  int rc = ref_run(&c, f, scratch, alpha, steps, rho, uh, F, &st);
  if (rc == 0) {
    double m, e, w;
    ref_reduce(&c, rho, uh, &m, &e, &w);
    if (!isfinite(m) || !isfinite(e) || !isfinite(w)) rc = -100;
    /* the same step with the force read from a field */
    rc = rc ? rc : ref_step_field(&c, f, scratch, alpha, F, rho, uh, NULL, &st);
  }
  free(rho), free(u), free(f), free(scratch), free(alpha), free(uh), free(F);
  return rc;
}

int main(void) {
  ref_set_threads(2);
  const int lats[3] = {REF_LATTICE_D2Q9, REF_LATTICE_D3Q19, REF_LATTICE_D3Q27};
  const int fss[4] = {REF_FS_NONE, REF_FS_GUO, REF_FS_SHANCHEN, REF_FS_EDM};
  const int forces[4] = {REF_FORCE_NONE, REF_FORCE_CONSTANT, REF_FORCE_SINUSOIDAL,
                         REF_FORCE_KOLMOGOROV};
  const int shapes[4][3] = {{5, 4, 3}, {1, 2, 4}, {2, 1, 1}, {3, 3, 2}};
  int cases = 0, bad = 0;
  for (int l = 0; l < 3; ++l)
    for (int col = 1; col <= 2; ++col)
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          for (int e = 0; e < 2; ++e)
            for (int s = 0; s < 4; ++s) {
              const int eq = e ? REF_EQ_EXACT : REF_EQ_MA3;
              if (eq == REF_EQ_EXACT && lats[l] == REF_LATTICE_D3Q19) continue;
              const int nz = lats[l] == REF_LATTICE_D2Q9 ? 1 : shapes[s][2];
              const int rc = run_case(lats[l], col, fss[a], forces[b], eq, shapes[s][0],
                                      shapes[s][1], nz, col == 2 ? 3e-2 : 1e-2, 3);
              ++cases;
              if (rc != 0) {
                fprintf(stderr, "case lattice %d collision %d fs %d force %d eq %d shape %d: rc %d\n",
                        lats[l], col, fss[a], forces[b], eq, s, rc);
                ++bad;
              }
            }
  /* entropic alpha in every branch: small deviation, capped, Newton, and a
   * cell whose Newton leaves its bounds */
  for (int l = 0; l < 3; ++l) {
    const int Q = ref_lattice_q(lats[l]);
    double fp[27], fn[27];
    const double scales[4] = {1e-5, 0.3, 0.02, 0.9};
    for (int k = 0; k < 4; ++k) {
      for (int i = 0; i < Q; ++i) {
        const double w = ref_lattice_w(lats[l], i);
        fn[i] = scales[k] * w * ((i * 7919 % 13) / 6.0 - 1.0);
        fp[i] = w + fn[i];
      }
      int br = -1, it = 0;
      const double al = ref_elbm_alpha(lats[l], fp, fn, 2.0, &br, &it);
      const double g = ref_entropic_g(lats[l], fp, fn, al);
      const double dg = ref_entropic_dg(lats[l], fp, fn, al);
      if (!isfinite(al) || br < 0 || br > 3 || isnan(g) || isnan(dg)) ++bad;
    }
  }
  /* initialisers */
  {
    ref_config c = {REF_LATTICE_D3Q27, 1, REF_FS_NONE, REF_FORCE_NONE, REF_EQ_MA3, 9, 10, 7,
                    0.6, {0, 0, 0}, {1, 1, 1}};
    double *u = xcalloc((size_t)3 * 630, sizeof(double));
    ref_init_taylor_green(&c, 0.05, u);
    if (ref_init_spectral_velocity(&c, 0.01, 3, 2.0, 7, u) != 0) ++bad;
    if (ref_init_spectral_velocity(&c, 0.01, 4, 2.0, 7, u) != -1) ++bad;  /* 2 kmax >= 7 */
    free(u);
    (void)ref_splitmix64(1234567, 0);
  }
  /* MLP reference: forward, loss, gradient, Adam on both network shapes */
  {
    const int shapes_mlp[2][3] = {{9, 32, 32}, {27, 64, 64}};
    for (int k = 0; k < 2; ++k) {
      const int nin = shapes_mlp[k][0], h1 = shapes_mlp[k][1], h2 = shapes_mlp[k][2];
      const int np = mlp_ref_param_count(nin, h1, h2), B = 17;
      double *p = xcalloc((size_t)np, sizeof(double)), *g = xcalloc((size_t)np, sizeof(double));
      double *m = xcalloc((size_t)np, sizeof(double)), *v = xcalloc((size_t)np, sizeof(double));
      double *x = xcalloc((size_t)B * nin, sizeof(double)), *t = xcalloc((size_t)B, sizeof(double));
      double *y = xcalloc((size_t)B, sizeof(double));
      for (int i = 0; i < np; ++i) p[i] = 0.1 * sin(1.7 * i);
      for (int i = 0; i < B * nin; ++i) x[i] = cos(0.3 * i);
      for (int i = 0; i < B; ++i) t[i] = 0.01 * i;
      mlp_ref_forward(p, nin, h1, h2, B, x, y);
      const double l0 = mlp_ref_loss(p, nin, h1, h2, B, x, t);
      const double l1 = mlp_ref_grad(p, nin, h1, h2, B, x, t, g);
      mlp_ref_adam(p, g, m, v, np, 1e-3, 0.9, 0.999, 1e-8, 1);
      if (!isfinite(l0) || !isfinite(l1)) ++bad;
      free(p), free(g), free(m), free(v), free(x), free(t), free(y);
    }
  }
  printf("oracle sanitizer driver: %d step cases, %d failures\n", cases, bad);
  return bad ? 1 : 0;
}

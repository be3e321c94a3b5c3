/*
 * mlp_ref.c -- FP64 reference of the alpha-predictor MLP used by the learned
 * entropic collision.  TEST INFRASTRUCTURE ONLY: it checks the HIP training
 * and inference kernels (gradients, losses, predictions); the product never
 * links it.  No reference counterpart exists (the reference solves alpha by
 * Newton only, include/metaLBM/Collision.h:328-349).
 *
 * Network: y = W3 relu(W2 relu(W1 x + b1) + b2) + b3, scalar output,
 * relu(s) = max(s, 0) (derivative taken as 0 at s = 0).
 * Parameter vector layout (shared with the HIP kernels):
 *   W1[h1][nin], b1[h1], W2[h2][h1], b2[h2], W3[h2], b3[1].
 * Loss: L = 1/(2B) sum_b (y_b - t_b)^2.
 *
 * Parity pinning: central finite differences of mlp_ref_loss against
 * mlp_ref_grad (tests/test_oracle_mlp.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

int mlp_ref_param_count(int nin, int h1, int h2) {
  return h1 * nin + h1 + h2 * h1 + h2 + h2 + 1;
}

/* Forward for one sample; a1/a2 receive the hidden activations. */
static double forward1(const double *p, int nin, int h1, int h2,
                       const double *x, double *a1, double *a2) {
  const double *W1 = p, *b1 = W1 + h1 * nin, *W2 = b1 + h1, *b2 = W2 + h2 * h1,
               *W3 = b2 + h2, *b3 = W3 + h2;
  for (int j = 0; j < h1; ++j) {
    double s = b1[j];
    for (int k = 0; k < nin; ++k) s += W1[j * nin + k] * x[k];
    a1[j] = s > 0.0 ? s : 0.0;
  }
  for (int j = 0; j < h2; ++j) {
    double s = b2[j];
    for (int k = 0; k < h1; ++k) s += W2[j * h1 + k] * a1[k];
    a2[j] = s > 0.0 ? s : 0.0;
  }
  double y = b3[0];
  for (int k = 0; k < h2; ++k) y += W3[k] * a2[k];
  return y;
}

void mlp_ref_forward(const double *p, int nin, int h1, int h2, int batch,
                     const double *x, double *y) {
  double *a1 = (double *)malloc(sizeof(double) * (size_t)(h1 + h2));
  double *a2 = a1 + h1;
  for (int b = 0; b < batch; ++b)
    y[b] = forward1(p, nin, h1, h2, x + (size_t)b * nin, a1, a2);
  free(a1);
}

double mlp_ref_loss(const double *p, int nin, int h1, int h2, int batch,
                    const double *x, const double *t) {
  double *a1 = (double *)malloc(sizeof(double) * (size_t)(h1 + h2));
  double *a2 = a1 + h1;
  double s = 0.0;
  for (int b = 0; b < batch; ++b) {
    double r = forward1(p, nin, h1, h2, x + (size_t)b * nin, a1, a2) - t[b];
    s += r * r;
  }
  free(a1);
  return s / (2.0 * batch);
}

/* Gradient of mlp_ref_loss with respect to every parameter; returns loss. */
double mlp_ref_grad(const double *p, int nin, int h1, int h2, int batch,
                    const double *x, const double *t, double *g) {
  const int np = mlp_ref_param_count(nin, h1, h2);
  const double *W2 = p + h1 * nin + h1, *W3 = W2 + h2 * h1 + h2;
  double *gW1 = g, *gb1 = gW1 + h1 * nin, *gW2 = gb1 + h1,
         *gb2 = gW2 + h2 * h1, *gW3 = gb2 + h2, *gb3 = gW3 + h2;
  memset(g, 0, sizeof(double) * (size_t)np);
  double *buf = (double *)malloc(sizeof(double) * (size_t)(2 * h1 + 2 * h2));
  double *a1 = buf, *a2 = a1 + h1, *d1 = a2 + h2, *d2 = d1 + h1;
  double loss = 0.0;
  for (int b = 0; b < batch; ++b) {
    const double *xb = x + (size_t)b * nin;
    double r = forward1(p, nin, h1, h2, xb, a1, a2) - t[b];
    loss += r * r;
    double dy = r / batch;
    gb3[0] += dy;
    for (int k = 0; k < h2; ++k) {
      gW3[k] += dy * a2[k];
      d2[k] = a2[k] > 0.0 ? dy * W3[k] : 0.0;
      gb2[k] += d2[k];
    }
    for (int j = 0; j < h1; ++j) d1[j] = 0.0;
    for (int k = 0; k < h2; ++k)
      for (int j = 0; j < h1; ++j) {
        gW2[k * h1 + j] += d2[k] * a1[j];
        d1[j] += W2[k * h1 + j] * d2[k];
      }
    for (int j = 0; j < h1; ++j) {
      d1[j] = a1[j] > 0.0 ? d1[j] : 0.0;
      gb1[j] += d1[j];
      for (int k = 0; k < nin; ++k) gW1[j * nin + k] += d1[j] * xb[k];
    }
  }
  free(buf);
  return loss / (2.0 * batch);
}

/* One Adam step (Kingma & Ba 2015) with bias correction at step t >= 1. */
void mlp_ref_adam(double *p, const double *g, double *m, double *v, int n,
                  double lr, double b1, double b2, double eps, int t) {
  const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
  for (int i = 0; i < n; ++i) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
  }
}

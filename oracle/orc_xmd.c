/*
 * ORACLE -- test infrastructure only (see orc_xmd.h).  Plain-C FP64
 * restatement of the reference's nocl per-step path; every function cites the
 * reference file:line it restates.  Not linked into, or called by, the
 * product library.
 */
#include "orc_xmd.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.141592653589793
#endif

#define TRI(n, k) ((size_t)((n) - 1) * (size_t)(n) / 2 + (size_t)((k) - 1))
#define BEAD(l, j, P) (((l) - 1) * (P) + (j) - 1)

/* ---------------------------------------------------------------- images */

/* md_simulation_x.c:858-865 */
int orc_periodic_image_count(double x, double L) {
  int c = (int)(fabs(x) / L);
  if (x > L) return -c;
  if (x < 0) return c + 1;
  return 0;
}

/* md_simulation_x.c:70-81: shift by L/2, fold into [0, L], shift back. */
double orc_minimum_image(double a, double L) {
  double s = a + L / 2;
  s += L * orc_periodic_image_count(s, L);
  return s - L / 2;
}

static double dist(const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < ORC_DIM; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
  return sqrt(s);
}

static double mi_dist(const double *a, const double *b, const double *L) {
  double s = 0;
  for (int i = 0; i < ORC_DIM; ++i) {
    double t = orc_minimum_image(a[i] - b[i], L[i]);
    s += t * t;
  }
  return sqrt(s);
}

/* ---------------------------------------------------------- interactions */

/* Reciprocal-space part shared by the Ewald force variants
 * (md_simulation_x.c:266-281, 303-318): cube of integer triples, origin
 * skipped. */
static void ewald_recip_force(const orc_params_t *p, const double *p1, const double *p2, double *f,
                              const double *L, double m) {
  int n = p->ewald_n;
  double V = L[0] * L[1] * L[2];
  double a2 = p->ewald_alpha * p->ewald_alpha;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double g[3] = {i / L[0], j / L[1], k / L[2]};
        double G = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        double dot = 2 * M_PI * (g[0] * (p1[0] - p2[0]) + g[1] * (p1[1] - p2[1]) + g[2] * (p1[2] - p2[2]));
        double w = p->coul_q * sin(dot) * pow(G, -2) * exp(-M_PI * M_PI * G * G / a2) / (V * M_PI);
        for (int d = 0; d < 3; ++d) f[d] += 2 * M_PI * g[d] * w / m;
      }
}

/* Real-space image cube (md_simulation_x.c:282-292); skip_min drops the
 * minimum image as the R split does (md_simulation_x.c:319-333). */
static void ewald_real_force(const orc_params_t *p, const double *p1, const double *p2, double *f,
                             const double *L, double m, int skip_min) {
  int n = p->ewald_n;
  double al = p->ewald_alpha;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        double r[3] = {p2[0] + i * L[0], p2[1] + j * L[1], p2[2] + k * L[2]};
        if (skip_min) {
          double dis[3];
          for (int d = 0; d < 3; ++d)
            dis[d] = fabs(r[d] - p1[d]) - fabs(orc_minimum_image(r[d] - p1[d], L[d]));
          if (fabs(dis[0]) < 1e-4 && fabs(dis[1]) < 1e-4 && fabs(dis[2]) < 1e-4) continue;
        }
        double R = dist(p1, r);
        double w = p->coul_q * 2 * al * exp(-al * al * R * R) / (sqrt(M_PI) * R * R);
        w += p->coul_q * erfc(al * R) / (R * R * R);
        for (int d = 0; d < 3; ++d) f[d] += (p1[d] - r[d]) * w / m;
      }
}

void orc_pair_force_one(int kind, const orc_params_t *p, const double *p1, const double *p2,
                        double *f, const double *L, double m) {
  double s2 = p->gauss_s * p->gauss_s;
  switch (kind) {
    case ORC_PAIR_GAUSSIAN: { /* md_simulation_x.c:166-172 */
      double d = dist(p1, p2);
      double w = (p->gauss_g / (M_PI * s2)) * exp(-d * d / s2);
      for (int i = 0; i < 3; ++i) f[i] += w * (2 * (p1[i] - p2[i])) / s2 / m;
      break;
    }
    case ORC_PAIR_PERIODIC_GAUSSIAN: { /* md_simulation_x.c:174-180 */
      double d = mi_dist(p1, p2, L);
      double w = (p->gauss_g / (M_PI * s2)) * exp(-d * d / s2);
      for (int i = 0; i < 3; ++i) f[i] += w * (2 * orc_minimum_image(p1[i] - p2[i], L[i])) / s2 / m;
      break;
    }
    case ORC_PAIR_EWALD: /* md_simulation_x.c:259-293 */
      ewald_recip_force(p, p1, p2, f, L, m);
      ewald_real_force(p, p1, p2, f, L, m, 0);
      break;
    case ORC_PAIR_EWALD_R: /* md_simulation_x.c:295-334 */
      ewald_recip_force(p, p1, p2, f, L, m);
      ewald_real_force(p, p1, p2, f, L, m, 1);
      break;
    case ORC_PAIR_EWALD_NI: { /* md_simulation_x.c:336-346 */
      double al = p->ewald_alpha;
      double R = mi_dist(p1, p2, L);
      double w = p->coul_q * 2 * al * exp(-al * al * R * R) / (sqrt(M_PI) * R * R);
      w += p->coul_q * erfc(al * R) / (R * R * R);
      for (int d = 0; d < 3; ++d) f[d] += orc_minimum_image(p1[d] - p2[d], L[d]) * w / m;
      break;
    }
    case ORC_PAIR_LJ: { /* md_simulation_x.c:116-122 */
      double r = mi_dist(p1, p2, L);
      double mult = 48 * pow(r, -14.0) - 24 * pow(r, -8.0);
      for (int i = 0; i < 3; ++i) f[i] += orc_minimum_image(p1[i] - p2[i], L[i]) * mult / m;
      break;
    }
    case ORC_PAIR_HELIUM: { /* md_simulation_x.c:385-399 */
      const double *h = p->he; /* eps, A, alpha, C6, C8, C10, D, rm */
      double r = mi_dist(p1, p2, L);
      double x = r / h[7];
      double F = x < h[6] ? exp(-pow(h[6] / x - 1, 2)) : 1;
      double mult = h[2] * h[1] * exp(-h[2] * x) - (6 * h[3] * pow(x, -7) + 8 * h[4] * pow(x, -9) + 10 * h[5] * pow(x, -11)) * F;
      if (x < h[6]) mult += (h[3] * pow(x, -6) + h[4] * pow(x, -8) + h[5] * pow(x, -10)) * F * 2 * h[6] * (h[6] / x - 1) / (x * x);
      for (int i = 0; i < 3; ++i) f[i] += h[0] * mult * orc_minimum_image(p1[i] - p2[i], L[i]) / (r * h[7] * m);
      break;
    }
    default:
      break;
  }
}

double orc_pair_energy_one(int kind, const orc_params_t *p, const double *p1, const double *p2,
                           const double *L) {
  double s2 = p->gauss_s * p->gauss_s;
  switch (kind) {
    case ORC_PAIR_GAUSSIAN: { /* md_statistics_x.c:340-344 */
      double d = dist(p1, p2);
      return (p->gauss_g / (M_PI * s2)) * exp(-d * d / s2);
    }
    case ORC_PAIR_PERIODIC_GAUSSIAN: { /* md_statistics_x.c:346-350 */
      double d = mi_dist(p1, p2, L);
      return (p->gauss_g / (M_PI * s2)) * exp(-d * d / s2);
    }
    case ORC_PAIR_EWALD: { /* md_statistics_x.c:364-396 */
      int n = p->ewald_n;
      double V = L[0] * L[1] * L[2];
      double al = p->ewald_alpha, a2 = al * al;
      double res = 0;
      for (int i = -n; i <= n; ++i)
        for (int j = -n; j <= n; ++j)
          for (int k = -n; k <= n; ++k) {
            if (i == 0 && j == 0 && k == 0) continue;
            double g[3] = {i / L[0], j / L[1], k / L[2]};
            double G = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            double dot = 2 * M_PI * (g[0] * (p1[0] - p2[0]) + g[1] * (p1[1] - p2[1]) + g[2] * (p1[2] - p2[2]));
            res += cos(dot) * pow(G, -2) * exp(-M_PI * M_PI * G * G / a2) / (V * M_PI);
          }
      res -= M_PI / (a2 * V);
      for (int i = -n; i <= n; ++i)
        for (int j = -n; j <= n; ++j)
          for (int k = -n; k <= n; ++k) {
            double r[3] = {p2[0] + i * L[0], p2[1] + j * L[1], p2[2] + k * L[2]};
            double R = dist(p1, r);
            res += erfc(al * R) / R;
          }
      return p->coul_q * res;
    }
    case ORC_PAIR_LJ: { /* md_statistics_x.c:317-320 */
      double r = mi_dist(p1, p2, L);
      return 4 * (pow(r, -12) - pow(r, -6));
    }
    case ORC_PAIR_HELIUM: { /* md_statistics_x.c:496-506 */
      const double *h = p->he;
      double x = mi_dist(p1, p2, L) / h[7];
      double F = x < h[6] ? exp(-pow(h[6] / x - 1, 2)) : 1;
      return h[0] * (h[1] * exp(-h[2] * x) - (h[3] * pow(x, -6) + h[4] * pow(x, -8) + h[5] * pow(x, -10)) * F);
    }
    default:
      return 0;
  }
}

/* md_simulation_x.c:223-257 */
double orc_madelung(const orc_params_t *p, const double *L) {
  int n = p->ewald_n;
  double V = L[0] * L[1] * L[2];
  double al = p->ewald_alpha, a2 = al * al;
  double res = 0;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double g[3] = {i / L[0], j / L[1], k / L[2]};
        double G = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        res += pow(G, -2) * exp(-M_PI * M_PI * G * G / a2) / (V * M_PI);
      }
  res -= M_PI / (a2 * V);
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double r[3] = {i * L[0], j * L[1], k * L[2]};
        double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        res += erfc(al * R) / R;
      }
  res -= 2 * al * pow(M_PI, -0.5);
  return p->coul_q * res;
}

/* ------------------------------------------------- per-slice PIMD forces */

/* md_pimd_x.c:1380-1389 with md_harmonic_trap_force_x (md_simulation_x.c:127-131) */
void orc_trap_force(const orc_params_t *p, int N, int P, orc_particle_t *b) {
  double w2 = p->trap_omega * p->trap_omega;
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j) {
      orc_particle_t *q = &b[BEAD(l, j, P)];
      for (int i = 0; i < 3; ++i) {
        double f = 0;
        f += -w2 * (q->x[i] - p->trap_center[i]);
        q->f[i] += f / P;
      }
    }
}

/* md_pimd_x.c:1660-1674: slice j, pairs l > l2, Newton's third law with the
 * mass of the higher-index bead. */
void orc_pair_force(int kind, const orc_params_t *p, int N, int P, const double *L, orc_particle_t *b) {
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l)
      for (int l2 = 1; l2 < l; ++l2) {
        orc_particle_t *a = &b[BEAD(l, j, P)], *c = &b[BEAD(l2, j, P)];
        double f[3] = {0, 0, 0};
        orc_pair_force_one(kind, p, a->x, c->x, f, L, a->m);
        for (int i = 0; i < 3; ++i) {
          a->f[i] += f[i] / P;
          c->f[i] -= f[i] / P;
        }
      }
}

/* md_pimd_x.c:3452-3467 (two species, same P): bead (l, j) of a against
 * bead (l2, j) of c; force on a divided by its mass (by the callback), the
 * reaction m_a f / P / m_c on c. */
void orc_pair_force_2(int kind, const orc_params_t *p, int NA, int NB, int P, const double *L, orc_particle_t *a,
                      orc_particle_t *c) {
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= NA; ++l)
      for (int l2 = 1; l2 <= NB; ++l2) {
        orc_particle_t *u = &a[BEAD(l, j, P)], *w = &c[BEAD(l2, j, P)];
        double f[3] = {0, 0, 0};
        orc_pair_force_one(kind, p, u->x, w->x, f, L, u->m);
        for (int i = 0; i < 3; ++i) {
          u->f[i] += f[i] / P;
          w->f[i] -= u->m * f[i] / P / w->m;
        }
      }
}

/* md_pimd_x.c:3614-3621 */
double orc_pair_energy_2(int kind, const orc_params_t *p, int NA, int NB, int P, const double *L,
                         const orc_particle_t *a, const orc_particle_t *c) {
  double e = 0;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= NA; ++l)
      for (int l2 = 1; l2 <= NB; ++l2)
        e += orc_pair_energy_one(kind, p, a[BEAD(l, j, P)].x, c[BEAD(l2, j, P)].x, L) / P;
  return e;
}

/* ------------------------------------------------------------ observables */

static double obs_dist(const double *a, const double *b, const double *L, int image) {
  return image ? mi_dist(a, b, L) : dist(a, b);
}

/* md_pimd_x.c:1944-1962: histogram of bead distances to center, += 1/P. */
void orc_density_distribution(int N, int P, const orc_particle_t *b, const double *L, int image,
                              const double *center, double rmax, int points, double *fx) {
  double incre = rmax / points;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      double r = obs_dist(b[BEAD(l, j, P)].x, center, L, image);
      int k = (int)(r / incre);
      if (k >= points) continue;
      fx[k] += 1.0 / P;
    }
}

/* md_pimd_x.c:3019-3036: bead 1 of l against bead pi of every j != l, += 1/norm. */
void orc_pair_correlation(int N, int P, const orc_particle_t *b, const double *L, int image, double rmax,
                          double norm, int pi, int points, double *fx) {
  double incre = rmax / points;
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= N; ++j) {
      if (l == j) continue;
      double r = obs_dist(b[BEAD(l, 1, P)].x, b[BEAD(j, pi, P)].x, L, image);
      int k = (int)(r / incre);
      if (k >= points) continue;
      fx[k] += 1.0 / norm;
    }
}

/* md_pimd_x.c:4967-4996 */
void orc_sk_structure(int N, int P, const orc_particle_t *b, double q0, double qincre, int pi, int points,
                      double *fx) {
  for (int j = 0; j < points; ++j) {
    double q = q0 + j * qincre, res = 0;
    for (int i = 0; i < 3; ++i) {
      double sum = 0, sum2 = 0;
      for (int l = 1; l <= N; ++l) sum += cos(q * b[BEAD(l, 1, P)].x[i]);
      for (int l = 1; l <= N; ++l) sum2 += cos(q * b[BEAD(l, pi, P)].x[i]);
      res += sum * sum2 / N;
      sum = sum2 = 0;
      for (int l = 1; l <= N; ++l) sum += sin(q * b[BEAD(l, 1, P)].x[i]);
      for (int l = 1; l <= N; ++l) sum2 += sin(q * b[BEAD(l, pi, P)].x[i]);
      res += sum * sum2 / N;
    }
    fx[j] += res / 3;
  }
}

/* 2-D bin index with the reference's clamping (md_pimd_x.c:2739-2744). */
static int clamp_bin(int i1, int i2, int n, int points) {
  int idx = i1 * n + i2;
  if (idx < 0) idx = 0;
  if (idx >= points) idx = points - 1;
  return idx;
}

/* md_pimd_x.c:2723-2747 (two = 0; a == b, every l, j incl. l == j) and
 * md_pimd_x.c:2841-2872 (two = 1; a against b, both orders): 2-D histogram of
 * (x_{l,1}, x_{j,pi}) along dimension 0, += 1/P per entry. */
void orc_itcf(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, double rmin, double rmax,
              int pi, int points, int two, double *fx) {
  int n = (int)sqrt(points);
  double incre = (rmax - rmin) / n;
  for (int l = 1; l <= NA; ++l)
    for (int j = 1; j <= NB; ++j) {
      int i1 = (int)((a[BEAD(l, 1, P)].x[0] - rmin) / incre);
      int i2 = (int)((b[BEAD(j, pi, P)].x[0] - rmin) / incre);
      fx[clamp_bin(i1, i2, n, points)] += 1.0 / P;
      if (two) fx[clamp_bin(i2, i1, n, points)] += 1.0 / P;
    }
}

/* md_pimd_x.c:3111-3138: bead 1 of every l of a against bead pi of every j of b. */
void orc_pair_correlation_2(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, const double *L,
                            int image, double rmax, double norm, int pi, int points, double *fx) {
  double incre = rmax / points;
  for (int l = 1; l <= NA; ++l)
    for (int j = 1; j <= NB; ++j) {
      double r = obs_dist(a[BEAD(l, 1, P)].x, b[BEAD(j, pi, P)].x, L, image);
      int k = (int)(r / incre);
      if (k >= points) continue;
      fx[k] += 1.0 / norm;
    }
}

/* md_pimd_x.c:3248-3280: in-plane distances of bead pairs of one slice whose
 * two beads both lie in the window box = {x0, x1, y0, y1}. */
void orc_pair_correlation_2d(int N, int P, const orc_particle_t *b, double rmax, double norm, const double *box,
                             int points, double *fx) {
  double incre = rmax / points;
  for (int pj = 1; pj <= P; ++pj)
    for (int l = 1; l <= N; ++l)
      for (int j = 1; j <= N; ++j) {
        if (l == j) continue;
        const double *u = b[BEAD(l, pj, P)].x, *w = b[BEAD(j, pj, P)].x;
        if (u[0] < box[0] || u[0] > box[1] || u[1] < box[2] || u[1] > box[3]) continue;
        if (w[0] < box[0] || w[0] > box[1] || w[1] < box[2] || w[1] > box[3]) continue;
        double r = sqrt(pow(u[0] - w[0], 2) + pow(u[1] - w[1], 2));
        int k = (int)(r / incre);
        if (k >= points) continue;
        fx[k] += 1.0 / norm;
      }
}

/* md_pimd_x.c:4869-4911 (same = 1, a == b, l != j) and 4913-4956 (same = 0):
 * 2-D histogram of |dx|, |dy| between bead 1 of l and bead pi of j. */
void orc_2d_pair_correlation(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b,
                             const double *L, int image, double rmax, double norm, int pi, int points, int same,
                             double *fx) {
  int n = (int)sqrt(points);
  double incre = rmax / n;
  for (int l = 1; l <= NA; ++l)
    for (int j = 1; j <= NB; ++j) {
      if (same && l == j) continue;
      const double *u = a[BEAD(l, 1, P)].x, *w = b[BEAD(j, pi, P)].x;
      double dx = u[0] - w[0], dy = u[1] - w[1];
      if (image) {
        dx = orc_minimum_image(dx, L[0]);
        dy = orc_minimum_image(dy, L[1]);
      }
      int i1 = (int)(fabs(dx) / incre), i2 = (int)(fabs(dy) / incre);
      fx[clamp_bin(i1, i2, n, points)] += 1.0 / norm;
    }
}

/* Sum of cos / sin (q . x) over bead slot `bead` of every particle. */
static void cs_sum(int N, int P, const orc_particle_t *b, int bead, const double *qv, int axis, double q,
                   double *c, double *s) {
  for (int l = 1; l <= N; ++l) {
    const double *x = b[BEAD(l, bead, P)].x;
    double t;
    if (axis >= 0) {
      t = q * x[axis];
    } else {
      t = 0;
      for (int i = 0; i < 3; ++i) t += qv[i] * x[i];
    }
    *c += cos(t);
    *s += sin(t);
  }
}

/* md_pimd_x.c:5050-5083 (full = 0: bead 1 of a against bead pi of b, / N_a)
 * and 5150-5203 (full = 1: a and b together on both sides, / (N_a + N_b)). */
void orc_sk_structure_2(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, double q0,
                        double qincre, int pi, int points, int full, double *fx) {
  for (int j = 0; j < points; ++j) {
    double q = q0 + j * qincre, res = 0;
    double norm = full ? NA + NB : NA;
    for (int i = 0; i < 3; ++i) {
      double c1 = 0, s1 = 0, c2 = 0, s2 = 0;
      cs_sum(NA, P, a, 1, NULL, i, q, &c1, &s1);
      if (full) cs_sum(NB, P, b, 1, NULL, i, q, &c1, &s1);
      if (full) cs_sum(NA, P, a, pi, NULL, i, q, &c2, &s2);
      cs_sum(NB, P, b, pi, NULL, i, q, &c2, &s2);
      res += c1 * c2 / norm;
      res += s1 * s2 / norm;
    }
    fx[j] += res / 3;
  }
}

/* md_pimd_x.c:5404-5450 (variant 0: a alone), 5266-5340 (variant 1: a and b
 * on both sides, / (N_a + N_b)) and 5342-5402 (variant 2: a and b against c,
 * / sqrt((N_a + N_b) N_c)); q vectors qis[3 j .. 3 j + 2], second bead
 * (pi - 1) % P + 1. */
void orc_sphere_sk(int variant, int NA, int NB, int NC, int P, const orc_particle_t *a, const orc_particle_t *b,
                   const orc_particle_t *c, int pi, const double *qis, int nq, double *fx) {
  int b2 = (pi - 1) % P + 1;
  for (int j = 0; j < nq; ++j) {
    const double *qv = qis + 3 * j;
    double c1 = 0, s1 = 0, c2 = 0, s2 = 0, norm;
    cs_sum(NA, P, a, 1, qv, -1, 0, &c1, &s1);
    if (variant >= 1) cs_sum(NB, P, b, 1, qv, -1, 0, &c1, &s1);
    if (variant <= 1) cs_sum(NA, P, a, b2, qv, -1, 0, &c2, &s2);
    if (variant == 1) cs_sum(NB, P, b, b2, qv, -1, 0, &c2, &s2);
    if (variant == 2) cs_sum(NC, P, c, b2, qv, -1, 0, &c2, &s2);
    if (variant == 0)
      norm = NA;
    else if (variant == 1)
      norm = NA + NB;
    else
      norm = sqrt((double)((NA + NB) * NC));
    fx[j] += c1 * c2 / norm + s1 * s2 / norm;
  }
}

/* md_pimd_x.c:1518-1522 with md_harmonic_trap_energy_x (md_statistics_x.c:322-325) */
double orc_trap_energy(const orc_params_t *p, int N, int P, const orc_particle_t *b) {
  double e = 0;
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j) {
      const orc_particle_t *q = &b[BEAD(l, j, P)];
      double d = dist(q->x, p->trap_center);
      e += 0.5 * q->m * p->trap_omega * p->trap_omega * d * d / P;
    }
  return e;
}

/* md_pimd_x.c:1813-1821 */
double orc_pair_energy(int kind, const orc_params_t *p, int N, int P, const double *L,
                       const orc_particle_t *b) {
  double e = 0;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l)
      for (int l2 = 1; l2 < l; ++l2)
        e += orc_pair_energy_one(kind, p, b[BEAD(l, j, P)].x, b[BEAD(l2, j, P)].x, L) / P;
  return e;
}

/* ------------------------------------------------------------- exchange */

static double spring(const orc_particle_t *a, const orc_particle_t *c, double omegaP, int image,
                     const double *L) {
  double d = image ? mi_dist(a->x, c->x, L) : dist(a->x, c->x);
  return 0.5 * a->m * omegaP * omegaP * d * d;
}

/* md_pimd_x.c:2103-2206 (worm, cut and pbc branches inactive in MD) */
void orc_fast_fill_ENk(int N, int P, double omegaP, int image, const double *L,
                       const orc_particle_t *b, double *ENk, double *Eint) {
  for (int l = 1; l <= N; ++l) {
    double e = 0;
    for (int j = 1; j < P; ++j) e += spring(&b[BEAD(l, j, P)], &b[BEAD(l, j + 1, P)], omegaP, image, L);
    Eint[l - 1] = e;
    ENk[TRI(l, 1)] = e + spring(&b[BEAD(l, P, P)], &b[BEAD(l, 1, P)], omegaP, image, L);
  }
  for (int l = 1; l <= N; ++l)
    for (int k = 2; k <= l; ++k) {
      double e = ENk[TRI(l, k - 1)] + Eint[l - k];
      e -= spring(&b[BEAD(l, P, P)], &b[BEAD(l - k + 2, 1, P)], omegaP, image, L);
      e += spring(&b[BEAD(l - k + 1, P, P)], &b[BEAD(l - k + 2, 1, P)], omegaP, image, L);
      e += spring(&b[BEAD(l, P, P)], &b[BEAD(l - k + 1, 1, P)], omegaP, image, L);
      ENk[TRI(l, k)] = e;
    }
}

/* md_pimd_x.c:521-549 */
static int next_bead(int l, int j, int n, int k, int P) {
  if (j == P) return (l == n) ? BEAD(n - k + 1, 1, P) : BEAD(l + 1, 1, P);
  return BEAD(l, j + 1, P);
}

static int prev_bead(int l, int j, int n, int k, int P) {
  if (j == 1) return (l == n - k + 1) ? BEAD(n, P, P) : BEAD(l - 1, P, P);
  return BEAD(l, j - 1, P);
}

/* md_pimd_x.c:569-591, 659-662: explicit walk around the k-cycle ending at n. */
void orc_fill_ENk_slow(int N, int P, double omegaP, int image, const double *L,
                       const orc_particle_t *b, double *ENk) {
  for (int n = 1; n <= N; ++n)
    for (int k = 1; k <= n; ++k) {
      double e = 0;
      for (int l = n - k + 1; l <= n; ++l)
        for (int j = 1; j <= P; ++j)
          e += spring(&b[BEAD(l, j, P)], &b[next_bead(l, j, n, k, P)], omegaP, image, L);
      ENk[TRI(n, k)] = e;
    }
}

/* md_pimd_x.c:681-686 */
static double xexp(double k, double E, double EE, double beta, double vi) {
  if (vi == 0.0) return exp(-beta * E + EE);
  return exp((k - 1) * log(vi) - beta * E + EE);
}

/* md_pimd_x.c:797-806 */
static double xminE(int n, double beta, double vi, const double *ENk, const double *VBN) {
  if (vi == 0.0) return beta * (ENk[TRI(n, 1)] + VBN[n - 1]);
  double r = 1e10;
  for (int k = 1; k <= n; ++k) {
    double t = -(k - 1) * log(vi) + beta * (ENk[TRI(n, k)] + VBN[n - k]);
    if (t < r) r = t;
  }
  return r;
}

/* md_pimd_x.c:935-950 */
void orc_fill_VB(int N, double beta, double vi, const double *ENk, double *VBN) {
  VBN[0] = 0.0;
  for (int n = 1; n <= N; ++n) {
    double t = xminE(n, beta, vi, ENk, VBN);
    double s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s += xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    VBN[n] = (t - log(s) + log(n)) / beta;
  }
}

/* md_pimd_x.c:2322-2333 */
static double fast_xminE(int u, int N, double beta, double vi, const double *ENk, const double *V) {
  if (vi == 0.0) return log(u) + beta * (ENk[TRI(u, 1)] + V[u]);
  double r = 1e10;
  for (int l = u; l <= N; ++l) {
    int k = l - u + 1;
    double t = -(k - 1) * log(vi) + log(l) + beta * (ENk[TRI(l, k)] + V[l]);
    if (t < r) r = t;
  }
  return r;
}

/* md_pimd_x.c:2337-2342 */
static double fast_xexp(double l, double k, double E, double EE, double beta, double vi) {
  if (vi == 0.0) return exp(-beta * E + EE - log(l));
  return exp((k - 1) * log(vi) - beta * E + EE - log(l));
}

/* md_pimd_x.c:2352-2365 (cut_Pj inactive): spring gradient of bead idx
 * towards its two ring neighbours. */
static void dspring(const orc_particle_t *b, int idx, int nxt, int prv, double omegaP, int image,
                    const double *L, double *g) {
  for (int i = 0; i < 3; ++i) {
    double a = b[idx].x[i] - b[nxt].x[i];
    double c = b[idx].x[i] - b[prv].x[i];
    if (image) {
      a = orc_minimum_image(a, L[i]);
      c = orc_minimum_image(c, L[i]);
    }
    g[i] = 0;
    g[i] += b[idx].m * omegaP * omegaP * a;
    g[i] += b[idx].m * omegaP * omegaP * c;
  }
}

/* md_pimd_x.c:2649-2716 */
void orc_fast_force_VB(int N, int P, double beta, double vi, double omegaP, int image,
                       const double *L, const double *ENk, const double *VBN,
                       orc_particle_t *b, double *Vt_out, double *G_out) {
  double *Vt = (double *)malloc(sizeof(double) * (N + 1));
  double *G = (double *)malloc(sizeof(double) * (size_t)N * N);
  Vt[N] = 0;
  for (int u = N; u >= 1; --u) {
    double t = fast_xminE(u, N, beta, vi, ENk, Vt);
    double s = 0;
    for (int l = u; l <= N; ++l) {
      int k = l - u + 1;
      if (vi == 0 && k != 1) continue;
      s += fast_xexp(l, k, ENk[TRI(l, k)] + Vt[l], t, beta, vi);
    }
    Vt[u - 1] = (t - log(s)) / beta;
  }
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= N; ++j) {
      double g;
      if (j > l + 1)
        g = 0;
      else if (j == l + 1)
        g = 1 - exp(-beta * (VBN[l] + Vt[l] - VBN[N]));
      else if (vi == 0 && l != j)
        g = 0;
      else if (vi == 0)
        g = exp(-beta * (VBN[j - 1] + ENk[TRI(l, l - j + 1)] + Vt[l] - VBN[N])) / l;
      else
        g = exp(-beta * (VBN[j - 1] + ENk[TRI(l, l - j + 1)] + Vt[l] - VBN[N]) + (l - j) * log(vi)) / l;
      G[(size_t)(l - 1) * N + (j - 1)] = g;
    }
  double gr[3];
  for (int l = 1; l <= N; ++l)
    for (int j = 2; j < P; ++j) {
      int idx = BEAD(l, j, P);
      dspring(b, idx, BEAD(l, j + 1, P), BEAD(l, j - 1, P), omegaP, image, L, gr);
      for (int i = 0; i < 3; ++i) b[idx].f[i] += -gr[i] / b[idx].m;
    }
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= N; ++j) {
      int idx = BEAD(l, 1, P);
      dspring(b, idx, BEAD(l, 2, P), BEAD(j, P, P), omegaP, image, L, gr);
      for (int i = 0; i < 3; ++i) b[idx].f[i] += -G[(size_t)(j - 1) * N + (l - 1)] * gr[i] / b[idx].m;
    }
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= N; ++j) {
      int idx = BEAD(l, P, P);
      dspring(b, idx, BEAD(l, P - 1, P), BEAD(j, 1, P), omegaP, image, L, gr);
      for (int i = 0; i < 3; ++i) b[idx].f[i] += -G[(size_t)(l - 1) * N + (j - 1)] * gr[i] / b[idx].m;
    }
  if (Vt_out) memcpy(Vt_out, Vt, sizeof(double) * (N + 1));
  if (G_out) memcpy(G_out, G, sizeof(double) * (size_t)N * N);
  free(Vt);
  free(G);
}

/* Extended-precision twin of orc_fast_fill_ENk + orc_fill_VB +
 * orc_fast_force_VB (the same formulas, md_pimd_x.c:814-953, 1969-2210,
 * 2212-2721), every intermediate in x86-64 80-bit long double (64-bit
 * significand): springs, cycle energies, the forward and backward
 * log-sum-exp recursions, the connection matrix G and the G-weighted spring
 * gradients.  On inputs where beta V_B ~ 1e4..1e5 the G entries 1 - exp(-x)
 * and exp(...) carry the absolute round-off of x in double precision
 * (~1e-11), which both the reference and any double-precision code inherit;
 * here that round-off is ~2^11 times smaller, so the result is the truth the
 * double-precision codes are measured against.  Output: the exchange and
 * ring-spring forces (accelerations) of every bead, rounded to double, and
 * V_B^(0..N). */
typedef long double ld;

static ld ld_comp(ld a, ld Lx, int image) { return image ? a - Lx * roundl(a / Lx) : a; }

static ld ld_spring(const orc_particle_t *a, const orc_particle_t *c, ld om, int image, const double *L) {
  ld d2 = 0;
  for (int i = 0; i < 3; ++i) {
    ld d = ld_comp((ld)a->x[i] - (ld)c->x[i], (ld)L[i], image);
    d2 += d * d;
  }
  return 0.5L * (ld)a->m * om * om * d2;
}

void orc_exchange_forces_ld(int N, int P, double beta_d, double vi_d, double omegaP, int image, const double *L,
                            const orc_particle_t *b, double *f_out, double *VBN_out) {
  const ld beta = beta_d, vi = vi_d, om = omegaP;
  const ld lnvi = vi > 0 ? logl(vi) : 0;
  ld *E = (ld *)malloc(sizeof(ld) * (size_t)N * (N + 1) / 2);
  ld *Eint = (ld *)malloc(sizeof(ld) * N);
  ld *V = (ld *)malloc(sizeof(ld) * (N + 1));
  ld *Vt = (ld *)malloc(sizeof(ld) * (N + 1));
  ld *G = (ld *)malloc(sizeof(ld) * (size_t)N * N);
  for (int l = 1; l <= N; ++l) {
    ld e = 0;
    for (int j = 1; j < P; ++j) e += ld_spring(&b[BEAD(l, j, P)], &b[BEAD(l, j + 1, P)], om, image, L);
    Eint[l - 1] = e;
    E[TRI(l, 1)] = e + ld_spring(&b[BEAD(l, P, P)], &b[BEAD(l, 1, P)], om, image, L);
  }
  for (int l = 1; l <= N; ++l)
    for (int k = 2; k <= l; ++k)
      E[TRI(l, k)] = E[TRI(l, k - 1)] + Eint[l - k] -
                     ld_spring(&b[BEAD(l, P, P)], &b[BEAD(l - k + 2, 1, P)], om, image, L) +
                     ld_spring(&b[BEAD(l - k + 1, P, P)], &b[BEAD(l - k + 2, 1, P)], om, image, L) +
                     ld_spring(&b[BEAD(l, P, P)], &b[BEAD(l - k + 1, 1, P)], om, image, L);
  /* forward: V^(n) = -(1/beta) ln (1/n) sum_k xi^(k-1) exp(-beta (E_n^(k) + V^(n-k))) */
  V[0] = 0;
  for (int n = 1; n <= N; ++n) {
    ld t = 1e4000L;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      ld x = -(k - 1) * lnvi + beta * (E[TRI(n, k)] + V[n - k]);
      if (x < t) t = x;
    }
    ld s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s += expl((k - 1) * lnvi - beta * (E[TRI(n, k)] + V[n - k]) + t);
    }
    V[n] = (t - logl(s) + logl((ld)n)) / beta;
  }
  /* backward: Vt^(u-1) = -(1/beta) ln sum_{l>=u} xi^(l-u) exp(-beta (E_l^(l-u+1) + Vt^(l))) / l */
  Vt[N] = 0;
  for (int u = N; u >= 1; --u) {
    ld t = 1e4000L;
    for (int l = u; l <= N; ++l) {
      int k = l - u + 1;
      if (vi == 0 && k != 1) continue;
      ld x = -(k - 1) * lnvi + logl((ld)l) + beta * (E[TRI(l, k)] + Vt[l]);
      if (x < t) t = x;
    }
    ld s = 0;
    for (int l = u; l <= N; ++l) {
      int k = l - u + 1;
      if (vi == 0 && k != 1) continue;
      s += expl((k - 1) * lnvi - beta * (E[TRI(l, k)] + Vt[l]) + t - logl((ld)l));
    }
    Vt[u - 1] = (t - logl(s)) / beta;
  }
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= N; ++j) {
      ld g = 0;
      if (j == l + 1)
        g = -expm1l(-beta * (V[l] + Vt[l] - V[N]));
      else if (j <= l && !(vi == 0 && l != j))
        g = expl(-beta * (V[j - 1] + E[TRI(l, l - j + 1)] + Vt[l] - V[N]) + (l - j) * lnvi) / l;
      G[(size_t)(l - 1) * N + (j - 1)] = g;
    }
  const ld k2 = om * om;
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j) {
      const int idx = BEAD(l, j, P);
      const ld m = b[idx].m;
      ld f[3] = {0, 0, 0};
      if (j > 1 && j < P) {
        for (int i = 0; i < 3; ++i) {
          ld a = ld_comp((ld)b[idx].x[i] - (ld)b[BEAD(l, j + 1, P)].x[i], (ld)L[i], image);
          ld c = ld_comp((ld)b[idx].x[i] - (ld)b[BEAD(l, j - 1, P)].x[i], (ld)L[i], image);
          f[i] = -(m * k2 * (a + c)) / m;
        }
      } else {
        /* bead 1: partners bead P of every j' weighted by G(j', l); bead P:
         * bead 1 of every j' weighted by G(l, j') */
        const int inner = (j == 1) ? BEAD(l, 2, P) : BEAD(l, P - 1, P);
        for (int jp = 1; jp <= N; ++jp) {
          const ld g = (j == 1) ? G[(size_t)(jp - 1) * N + (l - 1)] : G[(size_t)(l - 1) * N + (jp - 1)];
          if (g == 0) continue;
          const int other = (j == 1) ? BEAD(jp, P, P) : BEAD(jp, 1, P);
          for (int i = 0; i < 3; ++i) {
            ld a = ld_comp((ld)b[idx].x[i] - (ld)b[inner].x[i], (ld)L[i], image);
            ld c = ld_comp((ld)b[idx].x[i] - (ld)b[other].x[i], (ld)L[i], image);
            f[i] += -g * (m * k2 * (a + c)) / m;
          }
        }
      }
      for (int i = 0; i < 3; ++i) f_out[3 * idx + i] = (double)f[i];
    }
  if (VBN_out)
    for (int n = 0; n <= N; ++n) VBN_out[n] = (double)V[n];
  free(E);
  free(Eint);
  free(V);
  free(Vt);
  free(G);
}

/* md_pimd_x.c:967-995: gradient of the k-cycle ending at n w.r.t. bead (l, j). */
static void slow_dENk(const orc_particle_t *b, int P, int n, int k, int l, int j, double omegaP,
                      int image, const double *L, double *g) {
  if (!(l >= n - k + 1 && l <= n)) {
    g[0] = g[1] = g[2] = 0;
    return;
  }
  dspring(b, BEAD(l, j, P), next_bead(l, j, n, k, P), prev_bead(l, j, n, k, P), omegaP, image, L, g);
}

/* md_pimd_x.c:1071-1104: forward-mode gradient of the V_B recursion. */
void orc_slow_force_VB(int N, int P, double beta, double vi, double omegaP, int image,
                       const double *L, const double *ENk, const double *VBN, orc_particle_t *b) {
  double *res = (double *)malloc(sizeof(double) * 3 * (N + 1));
  double *grad = (double *)malloc(sizeof(double) * 3 * N);
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j) {
      for (int i = 0; i < 3; ++i) res[i] = 0;
      for (int n = 1; n <= N; ++n) {
        double t = xminE(n, beta, vi, ENk, VBN);
        double s2 = 0;
        for (int k = 1; k <= n; ++k) {
          if (vi == 0 && k != 1) continue;
          s2 += xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
        }
        for (int k = 1; k <= n; ++k) slow_dENk(b, P, n, k, l, j, omegaP, image, L, &grad[3 * (k - 1)]);
        for (int i = 0; i < 3; ++i) {
          double s = 0;
          for (int k = 1; k <= n; ++k) {
            if (vi == 0 && k != 1) continue;
            s += (grad[3 * (k - 1) + i] + res[3 * (n - k) + i]) * xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
          }
          res[3 * n + i] = s / s2;
        }
      }
      int idx = BEAD(l, j, P);
      for (int i = 0; i < 3; ++i) b[idx].f[i] += -res[3 * N + i] / b[idx].m;
    }
  free(res);
  free(grad);
}

/* ------------------------------------------------------------ estimators */

/* md_pimd_x.c:1268-1291 */
double orc_VBN_energy(int N, int P, double beta, double vi, const double *ENk, const double *VBN) {
  double *res = (double *)malloc(sizeof(double) * (N + 1));
  res[0] = 0;
  for (int n = 1; n <= N; ++n) {
    double t = xminE(n, beta, vi, ENk, VBN);
    double s2 = 0, s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s2 += xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s += (res[n - k] - ENk[TRI(n, k)]) * xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    res[n] = s / s2;
  }
  double e = res[N];
  e += 3.0 * N * P / (2 * beta);
  free(res);
  return e;
}

/* md_pimd_x.c:4376-4420 with pbc = 0: 1/2 m w^2 (x_a - x_b).(x_a - x_c) */
static double e2(const orc_particle_t *b, int ia, int ib, int ic, double omegaP, int image,
                 const double *L) {
  double r = 0;
  for (int i = 0; i < 3; ++i) {
    double d = b[ia].x[i] - b[ib].x[i];
    double d2 = b[ia].x[i] - b[ic].x[i];
    if (image) {
      d = orc_minimum_image(d, L[i]);
      d2 = orc_minimum_image(d2, L[i]);
    }
    r += 0.5 * b[ia].m * omegaP * omegaP * d * d2;
  }
  return r;
}

/* md_pimd_x.c:4524-4553 (nocl orientation: cycles k >= 3 are walked with
 * reversed links). */
void orc_fast_fill_ENk2(int N, int P, double omegaP, int image, const double *L,
                        const orc_particle_t *b, double *ENk2) {
  for (int l = 1; l <= N; ++l)
    ENk2[TRI(l, 1)] = e2(b, BEAD(l, 1, P), BEAD(l, P, P), BEAD(l, 1, P), omegaP, image, L);
  for (int l = 1; l <= N; ++l)
    for (int k = 2; k <= l; ++k) {
      double e = ENk2[TRI(l, k - 1)];
      e -= e2(b, BEAD(l, 1, P), BEAD(l - k + 2, P, P), BEAD(l - k + 2, 1, P), omegaP, image, L);
      e += e2(b, BEAD(l - k + 1, 1, P), BEAD(l - k + 2, P, P), BEAD(l - k + 2, 1, P), omegaP, image, L);
      e += e2(b, BEAD(l, 1, P), BEAD(l - k + 1, P, P), BEAD(l - k + 1, 1, P), omegaP, image, L);
      ENk2[TRI(l, k)] = e;
    }
}

/* md_pimd_x.c:3959-3987, 4050-4053 (forward link orientation). */
void orc_fill_ENk2_slow(int N, int P, double omegaP, int image, const double *L,
                        const orc_particle_t *b, double *ENk2) {
  for (int n = 1; n <= N; ++n)
    for (int k = 1; k <= n; ++k) {
      double r = 0;
      for (int l = n - k + 1; l <= n; ++l)
        r += e2(b, next_bead(l, P, n, k, P), BEAD(l, P, P), BEAD(l, 1, P), omegaP, image, L);
      ENk2[TRI(n, k)] = r;
    }
}

/* md_pimd_x.c:4289-4345 (pbc = 0) */
double orc_vi_virial_energy(int N, int P, double beta, double vi, int image, const double *L,
                            const double *ENk, const double *ENk2, const double *VBN,
                            const orc_particle_t *b) {
  double *res = (double *)malloc(sizeof(double) * (N + 1));
  res[0] = 0;
  for (int n = 1; n <= N; ++n) {
    double t = xminE(n, beta, vi, ENk, VBN);
    double s2 = 0, s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s2 += xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s += (res[n - k] - ENk2[TRI(n, k)]) * xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    res[n] = s / s2;
  }
  double e = res[N];
  e += 3.0 * N / (2 * beta);
  free(res);
  double r2 = 0;
  for (int j = 2; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      const orc_particle_t *q = &b[BEAD(l, j, P)], *q1 = &b[BEAD(l, 1, P)];
      for (int i = 0; i < 3; ++i) {
        double d = q->x[i] - q1->x[i];
        if (image) d = orc_minimum_image(d, L[i]);
        r2 += -d * q->f[i] * q->m;
      }
    }
  return e + r2 / 2.0;
}

/* md_pimd_x.c:5452-5517 (pbc = 0): the virial recursion of the energy
 * estimator enters as 2 res[N] / 3 + N / beta, plus the bead-1-relative and
 * box-centre-relative virials, each / 3. */
double orc_vi_virial_pressure(int N, int P, double beta, double vi, int image, const double *L,
                              const double *ENk, const double *ENk2, const double *VBN,
                              const orc_particle_t *b) {
  double *res = (double *)malloc(sizeof(double) * (N + 1));
  res[0] = 0;
  for (int n = 1; n <= N; ++n) {
    double t = xminE(n, beta, vi, ENk, VBN);
    double s2 = 0, s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s2 += xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    for (int k = 1; k <= n; ++k) {
      if (vi == 0 && k != 1) continue;
      s += (res[n - k] - ENk2[TRI(n, k)]) * xexp(k, ENk[TRI(n, k)] + VBN[n - k], t, beta, vi);
    }
    res[n] = s / s2;
  }
  double e = 2.0 * res[N] / 3;
  e += N / beta;
  free(res);
  double r2 = 0;
  for (int j = 2; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      const orc_particle_t *q = &b[BEAD(l, j, P)], *q1 = &b[BEAD(l, 1, P)];
      for (int i = 0; i < 3; ++i) {
        double d = q->x[i] - q1->x[i];
        if (image) d = orc_minimum_image(d, L[i]);
        r2 += -d * q->f[i] * q->m;
      }
    }
  e += r2 / 3;
  r2 = 0;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      const orc_particle_t *q = &b[BEAD(l, j, P)];
      for (int i = 0; i < 3; ++i) r2 += (q->x[i] - L[i] / 2.0) * q->f[i] * q->m;
    }
  return e + r2 / 3;
}

/* md_pimd_x.c:3914-3920 */
double orc_virial_energy(int N, int P, const double *L, const orc_particle_t *b) {
  double r = 0;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      const orc_particle_t *q = &b[BEAD(l, j, P)];
      for (int i = 0; i < 3; ++i) r += -(q->x[i] - L[i] / 2) * q->f[i] * q->m;
    }
  return r / 2.0;
}

/* md_pimd_x.c:3937-3954 */
double orc_centroid_virial_energy(int N, int P, double beta, int fc, const orc_particle_t *b) {
  double *c = (double *)calloc((size_t)3 * N, sizeof(double));
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j)
      for (int i = 0; i < 3; ++i) c[3 * (l - 1) + i] += b[BEAD(l, j, P)].x[i] / P;
  double r = 0;
  for (int j = 1; j <= P; ++j)
    for (int l = 1; l <= N; ++l) {
      const orc_particle_t *q = &b[BEAD(l, j, P)];
      for (int i = 0; i < 3; ++i) r += -(q->x[i] - c[3 * (l - 1) + i]) * q->f[i] * q->m;
    }
  free(c);
  return r / 2.0 + (3.0 * N - fc / ((double)P)) / (2 * beta);
}

/* md_pimd_x.c:4652-4663 */
static double xminE2(int n, double beta, double vi2, const double *ENk, const double *V2) {
  if (vi2 == 0.0) return beta * (ENk[TRI(n, 1)] + V2[n - 1]);
  double r = 1e10;
  for (int k = 1; k <= n; ++k) {
    double t = -(k - 1) * log(vi2) + beta * (ENk[TRI(n, k)] + V2[n - k]);
    if (t < r) r = t;
  }
  return r;
}

/* md_pimd_x.c:4675-4701 */
double orc_vi_sign(int N, double beta, double vi2, const double *ENk, const double *VBN) {
  double *V2 = (double *)malloc(sizeof(double) * (N + 1));
  double *sg = (double *)malloc(sizeof(double) * (N + 1));
  V2[0] = 0;
  sg[0] = 1;
  double sv = vi2 < 0 ? -1 : 1;
  double a = fabs(vi2);
  for (int n = 1; n <= N; ++n) {
    double t = xminE2(n, beta, a, ENk, V2);
    double s = 0;
    for (int k = 1; k <= n; ++k) {
      if (vi2 == 0 && k != 1) continue;
      s += sg[n - k] * pow(sv, k - 1) * xexp(k, ENk[TRI(n, k)] + V2[n - k], t, beta, a);
    }
    sg[n] = s >= 0 ? 1 : -1;
    V2[n] = (t - log(fabs(s)) + log(n)) / beta;
  }
  double r = sg[N] * exp(-beta * V2[N] + beta * VBN[N]);
  free(V2);
  free(sg);
  return r;
}

/* ------------------------------------------------------------ integrator */

/* md_simulation_x.c:599-613 */
static void nhc_force(const orc_nhc_t *c, const orc_particle_t *b, int f0, int f2, double T, int M,
                      double *a) {
  double s = 0;
  for (int j = 0; j < f2; ++j) {
    int q = f0 + j, pi = q / 3, pj = q % 3;
    s += b[pi].m * b[pi].v[pj] * b[pi].v[pj];
  }
  a[0] = (s - c->f * T) / c->Q[0];
  for (int j = 1; j < M; ++j) a[j] = (c->Q[j - 1] * c->vtheta[j - 1] * c->vtheta[j - 1] - T) / c->Q[j];
}

static void kick(orc_particle_t *b, int f0, int f2, double h, double vt0) {
  for (int j = 0; j < f2; ++j) {
    int q = f0 + j, pi = q / 3, pj = q % 3;
    b[pi].v[pj] = b[pi].v[pj] * exp(-0.5 * h * vt0) + 0.5 * h * b[pi].f[pj] * exp(-0.25 * h * vt0);
  }
}

static void odd_chain(orc_nhc_t *c, const double *a, double h, int M) {
  int M2 = M / 2;
  for (int j = 1; j <= M2; ++j) {
    double up = (j == M2) ? 0 : c->vtheta[2 * j];
    c->vtheta[2 * j - 1] = c->vtheta[2 * j - 1] * exp(-0.5 * h * up) + 0.5 * h * a[2 * j - 1] * exp(-0.25 * h * up);
  }
}

/* md_simulation_x.c:674-705 */
void orc_nhc_vv3_1(int nb, int Nf, orc_nhc_t *nhc, orc_particle_t *b, double T, double h, int M) {
  double a[ORC_NHC_SLOTS];
  int f0 = 0, M2 = M / 2;
  for (int i = 0; i < Nf; ++i) {
    orc_nhc_t *c = &nhc[i];
    int f = c->f;
    int f2 = (i == Nf - 1) ? 3 * nb - f0 : f;
    nhc_force(c, b, f0, f2, T, M, a);
    kick(b, f0, f2, h, c->vtheta[0]);
    for (int j = 1; j <= M2; ++j) c->theta[2 * j - 2] = c->theta[2 * j - 2] + h * c->vtheta[2 * j - 2] / 2;
    odd_chain(c, a, h, M);
    for (int j = 0; j < f2; ++j) {
      int q = f0 + j, pi = q / 3, pj = q % 3;
      b[pi].x[pj] = b[pi].x[pj] + h * b[pi].v[pj];
    }
    for (int j = 1; j <= M2; ++j) c->theta[2 * j - 1] = c->theta[2 * j - 1] + h * c->vtheta[2 * j - 1];
    nhc_force(c, b, f0, f2, T, M, a);
    for (int j = 1; j <= M2; ++j)
      c->vtheta[2 * j - 2] = c->vtheta[2 * j - 2] * exp(-h * c->vtheta[2 * j - 1]) + h * a[2 * j - 2] * exp(-0.5 * h * c->vtheta[2 * j - 1]);
    f0 += f;
  }
}

/* md_simulation_x.c:761-789 */
void orc_nhc_vv3_2(int nb, int Nf, orc_nhc_t *nhc, orc_particle_t *b, double T, double h, int M) {
  double a[ORC_NHC_SLOTS];
  int f0 = 0, M2 = M / 2;
  for (int i = 0; i < Nf; ++i) {
    orc_nhc_t *c = &nhc[i];
    int f = c->f;
    int f2 = (i == Nf - 1) ? 3 * nb - f0 : f;
    kick(b, f0, f2, h, c->vtheta[0]);
    for (int j = 1; j <= M2; ++j) c->theta[2 * j - 2] = c->theta[2 * j - 2] + h * c->vtheta[2 * j - 2] / 2;
    nhc_force(c, b, f0, f2, T, M, a);
    odd_chain(c, a, h, M);
    f0 += f;
  }
}

/* md_simulation_x.c:794-856 (single wrap) */
void orc_periodic_boundary(int nb, const double *L, orc_particle_t *b) {
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) {
      if (b[i].x[d] < 0)
        b[i].x[d] += L[d];
      else if (b[i].x[d] > L[d])
        b[i].x[d] -= L[d];
    }
}

/* md_statistics_x.c:960-968 */
double orc_kinetic(int nb, const orc_particle_t *b) {
  double s = 0;
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) s += 0.5 * b[i].m * b[i].v[d] * b[i].v[d];
  return s;
}

/* md_statistics_x.c:970-983 */
double orc_nhc_kinetic(int Nf, const orc_nhc_t *nhc, double T, int M) {
  double s = 0;
  for (int i = 0; i < Nf; ++i)
    for (int j = 0; j < M; ++j) s += 0.5 * nhc[i].Q[j] * nhc[i].vtheta[j] * nhc[i].vtheta[j];
  for (int i = 0; i < Nf; ++i) {
    s += nhc[i].f * T * nhc[i].theta[0];
    for (int j = 1; j < M; ++j) s += T * nhc[i].theta[j];
  }
  return s;
}

/* ------------------------------------------------------------------- RNG */

/* md_util_x.c:671-693 */
void orc_set_seed(unsigned int seed) { srand(seed); }

double orc_random_uniform(double a, double b) {
  double r = ((double)rand()) / RAND_MAX;
  return a + (b - a) * r;
}

double orc_random_gaussian(void) {
  static double second;
  static int has = 0;
  if (has) {
    has = 0;
    return second;
  }
  has = 1;
  double r = sqrt(-2 * log(orc_random_uniform(0, 1)));
  double ang = 2 * M_PI * orc_random_uniform(0, 1);
  second = r * sin(ang);
  return r * cos(ang);
}

/* md_util_x.c:695-717 */
void orc_init_maxwell_vel(int nb, orc_particle_t *b, double T) {
  double mean[3] = {0, 0, 0};
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) {
      b[i].v[d] = orc_random_gaussian() * sqrt(T / b[i].m);
      mean[d] += b[i].m * b[i].v[d];
    }
  for (int d = 0; d < 3; ++d) mean[d] /= nb;
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) b[i].v[d] -= mean[d] / b[i].m;
}

/* md_pimd_x.c:551-567 */
void orc_init_uniform_pos(int N, int P, const double *center, const double *L, double fluc,
                          orc_particle_t *b) {
  for (int i = 0; i < N; ++i) {
    double pos[3];
    for (int k = 0; k < 3; ++k) pos[k] = orc_random_uniform(center[k] - L[k] / 2, center[k] + L[k] / 2);
    for (int j = 0; j < P; ++j)
      for (int k = 0; k < 3; ++k) b[BEAD(i + 1, j + 1, P)].x[k] = pos[k] + fluc * (orc_random_uniform(0, 2) - 1);
  }
}

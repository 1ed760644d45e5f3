// Host-side part of the C ABI: interaction parameters and callback symbols,
// minimum-image helpers, random numbers and seeded initial conditions,
// object allocation, text I/O and block statistics.
//
// None of this is per-step compute: the per-step path (forces, exchange,
// integrator, estimators) runs in the HIP kernels of the other translation
// units.  The interaction callbacks defined here are the symbols drivers pass
// by address (md_simulation_x.h:38-39); the engine recognises them by
// identity and never calls them per pair.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "xmd_internal.h"

extern "C" {

// ------------------------------------------------------------ parameters
// Globals of md_simulation_x.c:124-221.
md_num_t_x md_trap_frequency_x;
md_num_t_x md_trap_center_x[MD_DIMENSION_X];
md_num_t_x md_gaussian_strength_x;
md_num_t_x md_gaussian_range_x;
md_num_t_x md_coulomb_strength_x;
md_num_t_x md_coulomb_truc_x;
int md_coulomb_n_sum_x;

void md_set_harmonic_trap_frequency_x(md_num_t_x f) { md_trap_frequency_x = f; }
void md_set_harmonic_trap_center_x(md_num_t_x *center) {
  for (int i = 0; i < MD_DIMENSION_X; ++i) md_trap_center_x[i] = center[i];
}
void md_set_gaussian_force_strength_x(md_num_t_x g) { md_gaussian_strength_x = g; }
void md_set_guassian_force_range_x(md_num_t_x s) { md_gaussian_range_x = s; }
void md_set_coulomb_force_strength_x(md_num_t_x g) { md_coulomb_strength_x = g; }
void md_set_coulomb_truc_x(md_num_t_x k) { md_coulomb_truc_x = k; }
void md_set_coulomb_n_sum_x(int n) { md_coulomb_n_sum_x = n; }

// md_simulation_x.c:365-383
md_num_t_x md_he_eps_x, md_he_A_x, md_he_alpha_x, md_he_C6_x, md_he_C8_x, md_he_C10_x, md_he_D_x, md_he_rm_x;
void md_set_helium_parameters_x(md_num_t_x eps, md_num_t_x A, md_num_t_x alpha, md_num_t_x C6, md_num_t_x C8,
                                md_num_t_x C10, md_num_t_x D, md_num_t_x rm) {
  md_he_eps_x = eps;
  md_he_A_x = A;
  md_he_alpha_x = alpha;
  md_he_C6_x = C6;
  md_he_C8_x = C8;
  md_he_C10_x = C10;
  md_he_D_x = D;
  md_he_rm_x = rm;
}

// --------------------------------------------------------- image helpers
// md_simulation_x.c:858-865
int md_periodic_image_count_x(md_num_t_x x, md_num_t_x L) {
  int c = (int)(fabs(x) / L);
  if (x > L) return -c;
  if (x < 0) return c + 1;
  return 0;
}

// md_simulation_x.c:70-81
md_num_t_x md_minimum_image_x(md_num_t_x a, md_num_t_x L) {
  double s = a + L / 2;
  s += L * md_periodic_image_count_x(s, L);
  return s - L / 2;
}

md_num_t_x md_distance_x(md_num_t_x *p1, md_num_t_x *p2) {
  double s = 0;
  for (int i = 0; i < MD_DIMENSION_X; ++i) s += (p1[i] - p2[i]) * (p1[i] - p2[i]);
  return sqrt(s);
}

md_num_t_x md_minimum_image_distance_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L) {
  double s = 0;
  for (int i = 0; i < MD_DIMENSION_X; ++i) {
    double t = md_minimum_image_x(p1[i] - p2[i], L[i]);
    s += t * t;
  }
  return sqrt(s);
}

md_num_t_x md_minimum_image_distance_2_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, int *w1, int *w2) {
  double s = 0;
  for (int i = 0; i < MD_DIMENSION_X; ++i) {
    double t = (p1[i] + w1[i] * L[i]) - (p2[i] + w2[i] * L[i]);
    s += t * t;
  }
  return sqrt(s);
}

// ------------------------------------------------------ callback symbols
// Single-pair forms of the interactions.  Kept so that driver code that takes
// their addresses (or evaluates one pair on the host for diagnostics) links;
// formulas follow md_simulation_x.c:127-346 and md_statistics_x.c:322-396.

void md_harmonic_trap_force_x(md_num_t_x *p1, md_num_t_x *f, md_num_t_x *MD_UNUSED_X(L), md_num_t_x MD_UNUSED_X(m)) {
  for (int i = 0; i < MD_DIMENSION_X; ++i) f[i] += -md_trap_frequency_x * md_trap_frequency_x * (p1[i] - md_trap_center_x[i]);
}

void md_gaussian_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *MD_UNUSED_X(L), md_num_t_x m) {
  double s2 = md_gaussian_range_x * md_gaussian_range_x;
  double d = md_distance_x(p1, p2);
  double w = (md_gaussian_strength_x / (M_PI * s2)) * exp(-d * d / s2);
  for (int i = 0; i < MD_DIMENSION_X; ++i) f[i] += w * (2 * (p1[i] - p2[i])) / s2 / m;
}

void md_periodic_gaussian_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  double s2 = md_gaussian_range_x * md_gaussian_range_x;
  double d = md_minimum_image_distance_x(p1, p2, L);
  double w = (md_gaussian_strength_x / (M_PI * s2)) * exp(-d * d / s2);
  for (int i = 0; i < MD_DIMENSION_X; ++i) f[i] += w * (2 * md_minimum_image_x(p1[i] - p2[i], L[i])) / s2 / m;
}

// Lennard-Jones, minimum image (md_simulation_x.c:116-122)
void md_LJ_periodic_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  const double r = md_minimum_image_distance_x(p1, p2, L);
  const double w = 48 * pow(r, -14.0) - 24 * pow(r, -8.0);
  for (int i = 0; i < MD_DIMENSION_X; ++i) f[i] += md_minimum_image_x(p1[i] - p2[i], L[i]) * w / m;
}

// helium damping function and the scalar parts of force and energy (md_simulation_x.c:385-399)
static double he_damp(double x) {
  if (x < md_he_D_x) {
    const double u = md_he_D_x / x - 1;
    return exp(-u * u);
  }
  return 1.0;
}

void md_periodic_helium_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  const double r = md_minimum_image_distance_x(p1, p2, L);
  const double x = r / md_he_rm_x, F = he_damp(x);
  double w = md_he_alpha_x * md_he_A_x * exp(-md_he_alpha_x * x) -
             (6 * md_he_C6_x * pow(x, -7) + 8 * md_he_C8_x * pow(x, -9) + 10 * md_he_C10_x * pow(x, -11)) * F;
  if (x < md_he_D_x)
    w += (md_he_C6_x * pow(x, -6) + md_he_C8_x * pow(x, -8) + md_he_C10_x * pow(x, -10)) * F * 2 * md_he_D_x *
         (md_he_D_x / x - 1) / (x * x);
  for (int i = 0; i < MD_DIMENSION_X; ++i)
    f[i] += md_he_eps_x * w * md_minimum_image_x(p1[i] - p2[i], L[i]) / (r * md_he_rm_x * m);
}

static void ewald_pair(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m, int mode) {
  // mode 0: full, 1: all real-space images but the minimum one, 2: minimum image only
  const int n = md_coulomb_n_sum_x;
  const double q = md_coulomb_strength_x, al = md_coulomb_truc_x;
  const double V = L[0] * L[1] * L[2];
  if (mode != 2) {
    for (int i = -n; i <= n; ++i)
      for (int j = -n; j <= n; ++j)
        for (int k = -n; k <= n; ++k) {
          if (i == 0 && j == 0 && k == 0) continue;
          double g[3] = {i / L[0], j / L[1], k / L[2]};
          double G2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
          double dot = 2 * M_PI * (g[0] * (p1[0] - p2[0]) + g[1] * (p1[1] - p2[1]) + g[2] * (p1[2] - p2[2]));
          double w = q * sin(dot) / G2 * exp(-M_PI * M_PI * G2 / (al * al)) / (V * M_PI);
          for (int d = 0; d < 3; ++d) f[d] += 2 * M_PI * g[d] * w / m;
        }
  }
  auto real_term = [&](const double *r, double R) {
    double w = q * 2 * al * exp(-al * al * R * R) / (sqrt(M_PI) * R * R) + q * erfc(al * R) / (R * R * R);
    for (int d = 0; d < 3; ++d) f[d] += r[d] * w / m;
  };
  if (mode == 2) {
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = md_minimum_image_x(p1[d] - p2[d], L[d]);
    real_term(r, sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
    return;
  }
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        double r[3] = {p1[0] - (p2[0] + i * L[0]), p1[1] - (p2[1] + j * L[1]), p1[2] - (p2[2] + k * L[2])};
        if (mode == 1) {
          bool same = true;
          for (int d = 0; d < 3; ++d)
            same = same && fabs(fabs(r[d]) - fabs(md_minimum_image_x(-r[d], L[d]))) < 1e-4;
          if (same) continue;
        }
        real_term(r, sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
      }
}

void md_coulomb_3d_ewald_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  ewald_pair(p1, p2, f, L, m, 0);
}
void md_coulomb_3d_ewald_force_R_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  ewald_pair(p1, p2, f, L, m, 1);
}
void md_coulomb_3d_ewald_force_NI_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m) {
  ewald_pair(p1, p2, f, L, m, 2);
}

md_num_t_x md_harmonic_trap_energy_x(md_num_t_x *p1, md_num_t_x *MD_UNUSED_X(L), md_num_t_x m) {
  double d = md_distance_x(p1, md_trap_center_x);
  return 0.5 * m * md_trap_frequency_x * md_trap_frequency_x * d * d;
}

md_num_t_x md_gaussian_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *MD_UNUSED_X(L), md_num_t_x MD_UNUSED_X(m)) {
  double s2 = md_gaussian_range_x * md_gaussian_range_x;
  double d = md_distance_x(p1, p2);
  return (md_gaussian_strength_x / (M_PI * s2)) * exp(-d * d / s2);
}

md_num_t_x md_periodic_gaussian_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  double s2 = md_gaussian_range_x * md_gaussian_range_x;
  double d = md_minimum_image_distance_x(p1, p2, L);
  return (md_gaussian_strength_x / (M_PI * s2)) * exp(-d * d / s2);
}

// md_statistics_x.c:317-320
md_num_t_x md_LJ_periodic_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  const double r = md_minimum_image_distance_x(p1, p2, L);
  return 4 * (pow(r, -12.0) - pow(r, -6.0));
}

// md_statistics_x.c:496-506
md_num_t_x md_periodic_helium_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  const double x = md_minimum_image_distance_x(p1, p2, L) / md_he_rm_x;
  const double disp = md_he_C6_x * pow(x, -6) + md_he_C8_x * pow(x, -8) + md_he_C10_x * pow(x, -10);
  return md_he_eps_x * (md_he_A_x * exp(-md_he_alpha_x * x) - disp * he_damp(x));
}

// One pair's Ewald energy and its splits (md_statistics_x.c:364-495):
// mode 0 the whole sum, 1 (R) without the minimum image's real-space term,
// 2 (R2) R plus -erf(alpha r)/r of the minimum image, 3 (NI) the minimum
// image's erfc(alpha r)/r alone.  The minimum image is recognised as the
// reference does, per component within 1e-4 of the minimum-image distance.
static double ewald_energy_pair(const double *p1, const double *p2, const double *L, int mode) {
  const int n = md_coulomb_n_sum_x;
  const double al = md_coulomb_truc_x, V = L[0] * L[1] * L[2];
  double pa[3] = {p1[0], p1[1], p1[2]}, pb[3] = {p2[0], p2[1], p2[2]};
  if (mode == 3) return md_coulomb_strength_x * erfc(al * md_minimum_image_distance_x(pa, pb, (double *)L)) /
                        md_minimum_image_distance_x(pa, pb, (double *)L);
  double res = 0;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double g[3] = {i / L[0], j / L[1], k / L[2]};
        double G2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        double dot = 2 * M_PI * (g[0] * (p1[0] - p2[0]) + g[1] * (p1[1] - p2[1]) + g[2] * (p1[2] - p2[2]));
        res += cos(dot) / G2 * exp(-M_PI * M_PI * G2 / (al * al)) / (V * M_PI);
      }
  res -= M_PI / (al * al * V);
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        double r[3] = {p1[0] - (p2[0] + i * L[0]), p1[1] - (p2[1] + j * L[1]), p1[2] - (p2[2] + k * L[2])};
        if (mode != 0) {
          bool same = true;
          for (int d = 0; d < 3 && same; ++d)
            same = fabs(fabs(r[d]) - fabs(md_minimum_image_x(-r[d], L[d]))) < 1e-4;
          if (same) continue;
        }
        double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        res += erfc(al * R) / R;
      }
  if (mode == 2) {
    const double R = md_minimum_image_distance_x(pa, pb, (double *)L);
    res += -erf(al * R) / R;
  }
  return md_coulomb_strength_x * res;
}

md_num_t_x md_coulomb_3d_ewald_energy_R_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  return ewald_energy_pair(p1, p2, L, 1);
}
md_num_t_x md_coulomb_3d_ewald_energy_R2_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  return ewald_energy_pair(p1, p2, L, 2);
}
md_num_t_x md_coulomb_3d_ewald_energy_NI_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  return ewald_energy_pair(p1, p2, L, 3);
}

md_num_t_x md_coulomb_3d_ewald_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x MD_UNUSED_X(m)) {
  const int n = md_coulomb_n_sum_x;
  const double al = md_coulomb_truc_x, V = L[0] * L[1] * L[2];
  double res = 0;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double g[3] = {i / L[0], j / L[1], k / L[2]};
        double G2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        double dot = 2 * M_PI * (g[0] * (p1[0] - p2[0]) + g[1] * (p1[1] - p2[1]) + g[2] * (p1[2] - p2[2]));
        res += cos(dot) / G2 * exp(-M_PI * M_PI * G2 / (al * al)) / (V * M_PI);
      }
  res -= M_PI / (al * al * V);
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        double r[3] = {p1[0] - (p2[0] + i * L[0]), p1[1] - (p2[1] + j * L[1]), p1[2] - (p2[2] + k * L[2])};
        double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        res += erfc(al * R) / R;
      }
  return md_coulomb_strength_x * res;
}

// Madelung constant of the one-component plasma (md_simulation_x.c:223-257):
// a set-up constant that the driver adds as N*madelung/2 to the energy.
md_num_t_x md_calc_3d_ueg_madelung_x(md_num_t_x *L) {
  const int n = md_coulomb_n_sum_x;
  const double al = md_coulomb_truc_x, V = L[0] * L[1] * L[2];
  double res = 0;
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double g[3] = {i / L[0], j / L[1], k / L[2]};
        double G = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        res += pow(G, -2) * exp(-M_PI * M_PI * G * G / (al * al)) / (V * M_PI);
      }
  res -= M_PI / (al * al * V);
  for (int i = -n; i <= n; ++i)
    for (int j = -n; j <= n; ++j)
      for (int k = -n; k <= n; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        double r[3] = {i * L[0], j * L[1], k * L[2]};
        double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        res += erfc(al * R) / R;
      }
  res -= 2 * al * pow(M_PI, -0.5);
  return md_coulomb_strength_x * res;
}

// ---------------------------------------------------------------- random
// glibc rand()/RAND_MAX and Box-Muller with a one-value cache, consumed in
// the same order as md_util_x.c:671-693 so that seeded inputs match.  The
// library also keeps the seed, the number of rand() draws since it and the
// cache, so that a checkpoint can restore the generator exactly.
static unsigned g_rng_seed = 1;  // glibc's seed before any srand
static unsigned long long g_rng_draws = 0;
static double g_bm_cached = 0;
static int g_bm_have = 0;

void md_set_seed_x(unsigned int seed) {
  srand(seed);
  g_rng_seed = seed;
  g_rng_draws = 0;
}

md_num_t_x md_random_uniform_x(md_num_t_x a, md_num_t_x b) {
  double r = ((double)rand()) / RAND_MAX;
  ++g_rng_draws;
  return a + (b - a) * r;
}

md_num_t_x md_random_gaussian_x(void) {
  if (g_bm_have) {
    g_bm_have = 0;
    return g_bm_cached;
  }
  g_bm_have = 1;
  double r = sqrt(-2 * log(md_random_uniform_x(0, 1)));
  double ang = 2 * M_PI * md_random_uniform_x(0, 1);
  // glibc's sincos: the reference build (gcc -O2) fuses its sin/cos pair into
  // this call, whose last bit can differ from separate sin and cos.
  double sn, cs;
  sincos(ang, &sn, &cs);
  g_bm_cached = r * sn;
  return r * cs;
}

md_num_t_x md_get_quick_Q_x(md_num_t_x T, md_num_t_x omega) { return T * MD_kB_X / (omega * omega); }

// ------------------------------------------------------------ allocation
int md_pimd_mode_x = 0;
void md_set_pimd_mode_x(int pimd_mode) { md_pimd_mode_x = pimd_mode; }

simulation_box_t_x *md_alloc_box_x(int box_type) {
  if (box_type != MD_ND_RECT_BOX_X) return NULL;
  simulation_box_t_x *b = (simulation_box_t_x *)malloc(sizeof(simulation_box_t_x));
  if (!b) return NULL;
  b->type = box_type;
  b->box = calloc(1, sizeof(nd_rect_t_x));
  if (!b->box) {
    free(b);
    return NULL;
  }
  return b;
}

// Thermostat grouping of md_util_x.c:553-583: ceil((3N - fc)/f) chains of f
// consecutive degrees of freedom, the last one shorter.
md_simulation_t_x *md_alloc_simulation_x(int N, int f, int fc, int box_type) {
  if (N <= 0 || f <= 0) return NULL;
  md_simulation_t_x *s = (md_simulation_t_x *)calloc(1, sizeof(md_simulation_t_x));
  if (!s) return NULL;
  s->N = N;
  s->particles = (md_particle_t_x *)calloc((size_t)N, sizeof(md_particle_t_x));
  int dof = MD_DIMENSION_X * N - fc;
  s->fc = dof;
  s->Nf = dof / f + (dof % f != 0);
  s->fs = (int *)malloc(sizeof(int) * s->Nf);
  s->f0s = (int *)malloc(sizeof(int) * s->Nf);
  s->nhcs = (md_nhc_t_x *)calloc((size_t)s->Nf, sizeof(md_nhc_t_x));
  s->box = md_alloc_box_x(box_type);
  if (!s->particles || !s->fs || !s->f0s || !s->nhcs || !s->box) {
    free(s->particles);
    free(s->fs);
    free(s->f0s);
    free(s->nhcs);
    free(s);
    return NULL;
  }
  for (int i = 0, f0 = 0; i < s->Nf; ++i) {
    s->fs[i] = ((i + 1) * f > dof) ? dof - i * f : f;
    s->f0s[i] = f0;
    f0 += s->fs[i];
  }
  s->pcount_in = N * (N + 1) / 2;
  s->pcount_ex = N * (N - 1) / 2;
  if (!md_pimd_mode_x) {
    s->pair_in = (md_index_pair_t_x *)malloc(sizeof(md_index_pair_t_x) * s->pcount_in);
    s->pair_ex = (md_index_pair_t_x *)malloc(sizeof(md_index_pair_t_x) * (s->pcount_ex > 0 ? s->pcount_ex : 1));
    if (!s->pair_in || !s->pair_ex) return NULL;
    int a = 0, b = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j <= i; ++j) {
        s->pair_in[a].i = i;
        s->pair_in[a++].j = j;
        if (j < i) {
          s->pair_ex[b].i = i;
          s->pair_ex[b++].j = j;
        }
      }
  }
  return s;
}

int md_init_nd_rect_box_x(simulation_box_t_x *box, md_num_t_x *L) {
  if (box->type != MD_ND_RECT_BOX_X) return -1;
  nd_rect_t_x *r = (nd_rect_t_x *)box->box;
  for (int i = 0; i < MD_DIMENSION_X; ++i) r->L[i] = L[i];
  return 0;
}

// md_util_x.c:719-727
int md_init_particle_mass_x(md_simulation_t_x *sim, md_num_t_x m, int start, int end) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  for (int i = start; i < end; ++i) sim->particles[i].m = m;
  return 0;
}

// md_util_x.c:695-717: per bead and dimension a Gaussian draw, then the mean
// momentum per dimension is removed.
int md_init_maxwell_vel_x(md_simulation_t_x *sim, md_num_t_x T) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  double mean[MD_DIMENSION_X] = {0, 0, 0};
  for (int i = 0; i < sim->N; ++i)
    for (int d = 0; d < MD_DIMENSION_X; ++d) {
      sim->particles[i].v[d] = md_random_gaussian_x() * sqrt(MD_kB_X * T / sim->particles[i].m);
      mean[d] += sim->particles[i].m * sim->particles[i].v[d];
    }
  for (int d = 0; d < MD_DIMENSION_X; ++d) mean[d] /= sim->N;
  for (int i = 0; i < sim->N; ++i)
    for (int d = 0; d < MD_DIMENSION_X; ++d) sim->particles[i].v[d] -= mean[d] / sim->particles[i].m;
  return 0;
}

// md_util_x.c:805-826: theta = 0, vtheta = 1, Q0 = f*Q (or fp*Q), Qk = Q.
int md_init_nhc_x(md_simulation_t_x *sim, md_num_t_x Q, int start, int end, md_num_t_x fp) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  for (int i = start; i < end; ++i) {
    md_nhc_t_x *c = &sim->nhcs[i];
    c->f = sim->fs[i];
    for (int j = 0; j < MD_NHC_LENGTH_X; ++j) {
      c->theta[j] = 0;
      c->vtheta[j] = 1;
      c->Q[j] = (j == 0) ? (fp == 0.0 ? c->f * Q : fp * Q) : Q;
    }
  }
  return 0;
}

// md_util_x.c:729-803: four interleaved simple-cubic sublattices.
int md_init_particle_face_center_3d_lattice_pos_x(md_simulation_t_x *sim, md_num_t_x *center, md_num_t_x *length,
                                                  md_num_t_x fluc, md_num_t_x eps) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  double M2 = pow(sim->N / 4.0, 1.0 / 3);
  int M = (int)M2;
  if (M2 - M != 0) ++M;
  int count = 0;
  // sublattice s shifts every axis except axis (s-1) by half a cell (s = 0: no shift)
  for (int s = 0; s < 4; ++s)
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j)
        for (int k = 0; k < M; ++k) {
          if (count >= sim->N) break;
          int cell[3] = {i, j, k};
          for (int d = 0; d < 3; ++d) {
            double x = center[d] - length[d] / 2.0 + cell[d] * length[d] / M;
            if (s != 0 && d != s - 1) x += length[d] / M / 2.0;
            sim->particles[count].x[d] = x + eps + md_random_uniform_x(-fluc, fluc);
          }
          ++count;
        }
  return 0;
}

int md_simulation_copy_to_x(md_simulation_t_x *sim, md_simulation_t_x *sim2) {
  if (sim->N != sim2->N || sim->Nf != sim2->Nf) return -1;
  if (!md_simulation_sync_host_x(sim, 0) || !md_simulation_sync_host_x(sim2, 1)) return -1;
  for (int i = 0; i < sim->N; ++i) {
    for (int d = 0; d < MD_DIMENSION_X; ++d) {
      sim->particles[i].x[d] = sim2->particles[i].x[d];
      sim->particles[i].v[d] = sim2->particles[i].v[d];
    }
    sim->particles[i].m = sim2->particles[i].m;
  }
  for (int i = 0; i < sim->Nf; ++i) sim->nhcs[i] = sim2->nhcs[i];
  return 0;
}

// ------------------------------------------------------------ text I/O
// Formats of md_util_x.c:838-1011 ("%f", six decimals: lossy by design).
// The checkpoint pair below writes the same layout with "%a" (exact).
static int print_vec(FILE *out, md_simulation_t_x *sim, const char *tag, int which, const char *fmt = "%f ") {
  fprintf(out, "%s %d %d\n", tag, sim->N, MD_DIMENSION_X);
  if (!md_simulation_sync_host_x(sim, 1)) return -1;
  for (int i = 0; i < sim->N; ++i) {
    const double *v = which == 0 ? sim->particles[i].x : (which == 1 ? sim->particles[i].v : sim->particles[i].f);
    for (int d = 0; d < MD_DIMENSION_X; ++d) fprintf(out, fmt, v[d]);
    fprintf(out, "\n");
  }
  return ferror(out);
}
int md_fprint_particle_pos_x(FILE *out, md_simulation_t_x *sim) { return print_vec(out, sim, "x", 0); }
int md_fprint_particle_vel_x(FILE *out, md_simulation_t_x *sim) { return print_vec(out, sim, "v", 1); }
int md_fprint_particle_force_x(FILE *out, md_simulation_t_x *sim) { return print_vec(out, sim, "f", 2); }

static int print_nhcs(FILE *out, md_simulation_t_x *sim, const char *fmt) {
  fprintf(out, "nhc %d %d\n", sim->Nf, MD_NHC_LENGTH_X);
  if (!md_simulation_sync_host_x(sim, 1)) return -1;
  for (int i = 0; i < sim->Nf; ++i) {
    const md_nhc_t_x *c = &sim->nhcs[i];
    fprintf(out, "%d\n", c->f);
    for (int j = 0; j < MD_NHC_LENGTH_X; ++j) fprintf(out, fmt, c->theta[j]);
    fprintf(out, "\n");
    for (int j = 0; j < MD_NHC_LENGTH_X; ++j) fprintf(out, fmt, c->vtheta[j]);
    fprintf(out, "\n");
    for (int j = 0; j < MD_NHC_LENGTH_X; ++j) fprintf(out, fmt, c->Q[j]);
    fprintf(out, "\n");
  }
  return ferror(out);
}

int md_fprint_nhcs_x(FILE *out, md_simulation_t_x *sim) { return print_nhcs(out, sim, "%f "); }

int md_fprint_sim_box_x(FILE *out, simulation_box_t_x *box) {
  fprintf(out, "box %d %d\n", box->type, MD_DIMENSION_X);
  if (box->type == MD_ND_RECT_BOX_X) {
    nd_rect_t_x *r = (nd_rect_t_x *)box->box;
    for (int i = 0; i < MD_DIMENSION_X; ++i) fprintf(out, "%f ", r->L[i]);
    fprintf(out, "\n");
  }
  return ferror(out);
}

static int read_header(FILE *in, int expect_n, int expect_d) {
  char name[32];
  int n, d;
  if (fscanf(in, "%31s", name) != 1) return -1;
  if (fscanf(in, "%d", &n) != 1 || n != expect_n) return -1;
  if (fscanf(in, "%d", &d) != 1 || d != expect_d) return -1;
  return 0;
}

static int read_vec(FILE *in, md_simulation_t_x *sim, int which) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  if (read_header(in, sim->N, MD_DIMENSION_X)) return -1;
  for (int i = 0; i < sim->N; ++i)
    for (int d = 0; d < MD_DIMENSION_X; ++d) {
      double v;
      if (fscanf(in, "%lf", &v) != 1) return -1;
      (which == 0 ? sim->particles[i].x : (which == 1 ? sim->particles[i].v : sim->particles[i].f))[d] = v;
    }
  return ferror(in);
}
int md_fread_particle_pos_x(FILE *in, md_simulation_t_x *sim) { return read_vec(in, sim, 0); }
int md_fread_particle_vel_x(FILE *in, md_simulation_t_x *sim) { return read_vec(in, sim, 1); }

int md_fread_nhcs_x(FILE *in, md_simulation_t_x *sim) {
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  if (read_header(in, sim->Nf, MD_NHC_LENGTH_X)) return -1;
  for (int i = 0; i < sim->Nf; ++i) {
    int f;
    if (fscanf(in, "%d", &f) != 1) return -1;
    sim->nhcs[i].f = f;
    double *arr[3] = {sim->nhcs[i].theta, sim->nhcs[i].vtheta, sim->nhcs[i].Q};
    for (int a = 0; a < 3; ++a)
      for (int j = 0; j < MD_NHC_LENGTH_X; ++j) {
        double v;
        if (fscanf(in, "%lf", &v) != 1) return -1;
        arr[a][j] = v;
      }
  }
  return ferror(in);
}

// Lossless checkpoint: the x, v, f and nhc blocks of the reference's text
// layout (md_util_x.c:838-915), every number a C99 hexadecimal float.  The
// reference's readers (fscanf "%lf") parse it too.
// A last line "rng <seed> <draws> <cache flag> <cache>" holds the library's
// random-number state; the reader restores it by reseeding and replaying the
// draws (checkpoints without the line leave the generator as it is).
int md_checkpoint_write_x(FILE *out, md_simulation_t_x *sim) {
  if (!out || !sim) return -1;
  if (print_vec(out, sim, "x", 0, "%a ") || print_vec(out, sim, "v", 1, "%a ") || print_vec(out, sim, "f", 2, "%a ") ||
      print_nhcs(out, sim, "%a "))
    return -1;
  fprintf(out, "rng %u %llu %d %a\n", g_rng_seed, g_rng_draws, g_bm_have, g_bm_cached);
  return fflush(out) || ferror(out) ? -1 : 0;
}

int md_checkpoint_read_x(FILE *in, md_simulation_t_x *sim) {
  if (!in || !sim) return -1;
  if (read_vec(in, sim, 0) || read_vec(in, sim, 1) || read_vec(in, sim, 2) || md_fread_nhcs_x(in, sim)) return -1;
  char tag[8] = {0};
  if (fscanf(in, "%7s", tag) != 1) return 0;  // no random-number state recorded
  unsigned seed;
  unsigned long long draws;
  int have;
  double cached;
  if (strcmp(tag, "rng") != 0 || fscanf(in, "%u %llu %d %lf", &seed, &draws, &have, &cached) != 4) return -1;
  md_set_seed_x(seed);
  for (unsigned long long i = 0; i < draws; ++i) (void)rand();
  g_rng_draws = draws;
  g_bm_have = have;
  g_bm_cached = cached;
  return 0;
}

// ------------------------------------------------------- key/value input
md_attr_pair_t_x *md_alloc_attr_pair_x(void) {
  return (md_attr_pair_t_x *)calloc(1, sizeof(md_attr_pair_t_x));
}

// md_util_x.c:95-118: whitespace separated "name value" pairs until EOF.
md_attr_pair_t_x *md_read_attr_pair_x(FILE *in, md_attr_pair_t_x *attr) {
  md_attr_pair_t_x *r = attr ? attr : md_alloc_attr_pair_x();
  if (!r) return NULL;
  char name[MD_MAX_NAME_LENGTH_X + 1];
  double value;
  while (fscanf(in, "%" MD_STRINGIFY_X(MD_MAX_NAME_LENGTH_X) "s", name) == 1) {
    if (strlen(name) == 0) break;
    if (fscanf(in, "%lf", &value) != 1) return NULL;
    r->N++;
    r->names = (md_name_t_x *)realloc(r->names, sizeof(md_name_t_x) * r->N);
    r->values = (md_num_t_x *)realloc(r->values, sizeof(md_num_t_x) * r->N);
    if (!r->names || !r->values) return NULL;
    strcpy(r->names[r->N - 1], name);
    r->values[r->N - 1] = value;
    name[0] = '\0';
  }
  return r;
}

md_num_t_x md_get_attr_value_x(md_attr_pair_t_x *attr, const char *name, int *found) {
  if (found) *found = 0;
  for (int i = 0; i < attr->N; ++i)
    if (strcmp(name, attr->names[i]) == 0) {
      if (found) *found = 1;
      return attr->values[i];
    }
  return 0;
}

// ------------------------------------------------------------ statistics
// md_statistics_x.c:203-315: samples accumulate into the open block; every
// MD_MAX_STATS_COUNT samples the block is closed (divided by its count).
md_stats_t_x *md_alloc_stats_x(int points) {
  md_stats_t_x *s = (md_stats_t_x *)calloc(1, sizeof(md_stats_t_x));
  if (!s) return NULL;
  s->N = 1;
  s->es = (md_num_t_x *)calloc(1, sizeof(md_num_t_x));
  s->Ts = (md_num_t_x *)calloc(1, sizeof(md_num_t_x));
  s->points = points;
  s->fx = (md_num_t_x *)calloc(points > 0 ? points : 1, sizeof(md_num_t_x));
  s->fxs = (md_num_t_x **)malloc(sizeof(md_num_t_x *));
  if (!s->es || !s->Ts || !s->fx || !s->fxs) return NULL;
  s->fxs[0] = (md_num_t_x *)calloc(points > 0 ? points : 1, sizeof(md_num_t_x));
  if (!s->fxs[0]) return NULL;
  return s;
}

md_stats_t_x *md_stats_sync_host_x(md_stats_t_x *stats) { return stats; }

// Spherical S(q) grid (md_statistics_x.c:104-161): q = (qx, qy, qz) on the
// cube q0, q0 + qincre, ... <= qmax per axis (the axis values accumulate by
// repeated addition, as in the reference's loops), |q| < 1e-4 dropped; the
// first vector of each |q| shell (shells closer than 1e-4 merged) is its
// representative, and the shells are sorted by |q|.
md_sk_sphere_t_x *md_alloc_sk_sphere_3d_x(md_num_t_x q0, md_num_t_x qincre, md_num_t_x qmax) {
  if (!(qincre > 0)) return nullptr;
  std::vector<double> axis;
  for (double q = q0; q <= qmax; q += qincre) axis.push_back(q);
  std::vector<double> qs, qis;
  for (double qx : axis)
    for (double qy : axis)
      for (double qz : axis) {
        const double q = sqrt(qx * qx + qy * qy + qz * qz);
        if (fabs(q) < 1e-4) continue;
        bool found = false;
        for (double v : qs)
          if (fabs(v - q) < 1e-4) {
            found = true;
            break;
          }
        if (found) continue;
        qs.push_back(q);
        qis.insert(qis.end(), {qx, qy, qz});
      }
  std::vector<int> ord(qs.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return qs[a] < qs[b]; });
  md_sk_sphere_t_x *res = (md_sk_sphere_t_x *)malloc(sizeof(md_sk_sphere_t_x));
  if (!res) return nullptr;
  res->N = (int)qs.size();
  res->qs = (md_num_t_x *)malloc(sizeof(md_num_t_x) * (qs.size() + 1));
  res->qis = (md_num_t_x *)malloc(sizeof(md_num_t_x) * (qis.size() + 3));
  if (!res->qs || !res->qis) {
    free(res->qs);
    free(res->qis);
    free(res);
    return nullptr;
  }
  for (int i = 0; i < res->N; ++i) {
    res->qs[i] = qs[ord[i]];
    for (int d = 0; d < 3; ++d) res->qis[3 * i + d] = qis[3 * ord[i] + d];
  }
  return res;
}

md_stats_t_x *md_update_stats_x(md_stats_t_x *s) {
  s->count++;
  if (s->count >= MD_MAX_STATS_COUNT) {
    s->es[s->N - 1] /= MD_MAX_STATS_COUNT;
    s->Ts[s->N - 1] /= MD_MAX_STATS_COUNT;
    for (int i = 0; i < s->points; ++i) s->fxs[s->N - 1][i] /= MD_MAX_STATS_COUNT;
    s->N++;
    s->count = 0;
    s->es = (md_num_t_x *)realloc(s->es, sizeof(md_num_t_x) * s->N);
    s->Ts = (md_num_t_x *)realloc(s->Ts, sizeof(md_num_t_x) * s->N);
    s->fxs = (md_num_t_x **)realloc(s->fxs, sizeof(md_num_t_x *) * s->N);
    if (!s->es || !s->Ts || !s->fxs) return NULL;
    s->es[s->N - 1] = 0;
    s->Ts[s->N - 1] = 0;
    s->fxs[s->N - 1] = (md_num_t_x *)calloc(s->points > 0 ? s->points : 1, sizeof(md_num_t_x));
    if (!s->fxs[s->N - 1]) return NULL;
  }
  return s;
}

md_stats_t_x *md_finalize_stats_x(md_stats_t_x *s) {
  s->e = 0;
  s->T = 0;
  for (int i = 0; i < s->points; ++i) s->fx[i] = 0;
  if (s->count != 0) {
    s->es[s->N - 1] /= s->count;
    s->Ts[s->N - 1] /= s->count;
    for (int i = 0; i < s->points; ++i) s->fxs[s->N - 1][i] /= s->count;
  } else {
    s->N--;
  }
  for (int b = 0; b < s->N; ++b) {
    s->e += s->es[b];
    s->T += s->Ts[b];
    for (int j = 0; j < s->points; ++j) s->fx[j] += s->fxs[b][j];
  }
  s->e /= s->N;
  s->T /= s->N;
  for (int j = 0; j < s->points; ++j) s->fx[j] /= s->N;
  return s;
}

// md_pimd_x.c:4558-4650 (scalar helpers)
int md_reset_stats_x(md_stats_t_x *stats, md_num_t_x e) {
  stats->es[0] = e;
  return 0;
}
int md_stats_add_to_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  stats->es[stats->N - 1] += stats2->es[0];
  return 0;
}
int md_stats_copy_to_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  stats->es[0] = stats2->es[0];
  return 0;
}
int md_stats_mult_to_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  stats->es[0] *= stats2->es[0];
  return 0;
}
// Histogram forms of the same helpers (md_pimd_x.c:4565-4650).
int md_reset_stats_f_x(md_stats_t_x *stats, md_num_t_x e) {
  for (int i = 0; i < stats->points; ++i) stats->fxs[0][i] = e;
  return 0;
}
int md_stats_add_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  if (stats->points != stats2->points) return -1;
  for (int i = 0; i < stats->points; ++i) stats->fxs[stats->N - 1][i] += stats2->fxs[0][i];
  return 0;
}
int md_stats_copy_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  if (stats->points != stats2->points) return -1;
  for (int i = 0; i < stats->points; ++i) stats->fxs[0][i] = stats2->fxs[0][i];
  return 0;
}
int md_stats_mult_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2) {
  for (int i = 0; i < stats->points; ++i) stats->fxs[0][i] *= stats2->es[0];
  return 0;
}

// ----------------------------------------------------- PIMD index helpers
// md_pimd_x.c:521-549: neighbours along the k-cycle that ends at particle n.
int md_pimd_next_index_x(int l, int j, int N2, int k, int P) {
  if (j == P) return (l == N2) ? MD_PIMD_INDEX_X(N2 - k + 1, 1, P) : MD_PIMD_INDEX_X(l + 1, 1, P);
  return MD_PIMD_INDEX_X(l, j + 1, P);
}

int md_pimd_prev_index_x(int l, int j, int N2, int k, int P) {
  if (j == 1) return (l == N2 - k + 1) ? MD_PIMD_INDEX_X(N2, P, P) : MD_PIMD_INDEX_X(l - 1, P, P);
  return MD_PIMD_INDEX_X(l, j - 1, P);
}

// md_pimd_x.c:681-686, 2337-2342: single weight of the recursions.
md_num_t_x md_pimd_xexp_x(md_num_t_x k, md_num_t_x E, md_num_t_x EE, md_num_t_x beta, md_num_t_x vi) {
  if (vi == 0.0) return exp(-beta * E + EE);
  return exp((k - 1) * log(vi) - beta * E + EE);
}

md_num_t_x md_pimd_fast_xexp_x(md_num_t_x l, md_num_t_x k, md_num_t_x E, md_num_t_x EE, md_num_t_x beta, md_num_t_x vi) {
  if (vi == 0.0) return exp(-beta * E + EE - log(l));
  return exp((k - 1) * log(vi) - beta * E + EE - log(l));
}

// md_pimd_x.c:551-567: one uniform centre per particle, beads scattered by
// +-fluc around it.
int md_pimd_init_particle_uniform_pos_x(md_pimd_t_x *pimd, md_num_t_x *center, md_num_t_x *L, md_num_t_x fluc) {
  md_simulation_t_x *sim = pimd->sim;
  if (!md_simulation_sync_host_x(sim, 0)) return -1;
  for (int i = 0; i < pimd->N; ++i) {
    double pos[3];
    for (int k = 0; k < 3; ++k) pos[k] = md_random_uniform_x(center[k] - L[k] / 2, center[k] + L[k] / 2);
    for (int j = 0; j < pimd->P; ++j) {
      md_particle_t_x *b = &sim->particles[MD_PIMD_INDEX_X(i + 1, j + 1, pimd->P)];
      for (int k = 0; k < 3; ++k) b->x[k] = pos[k] + fluc * (md_random_uniform_x(0, 2) - 1);
    }
  }
  return 0;
}

static int print_tri(FILE *out, const char *tag, md_num_t_x **T, int N) {
  fprintf(out, "%s %d\n", tag, N);
  for (int l = 1; l <= N; ++l) {
    for (int j = 1; j <= l; ++j) fprintf(out, "%f ", T[l - 1][j - 1]);
    fprintf(out, "\n");
  }
  return ferror(out);
}

int md_pimd_fprint_ENk_x(FILE *out, md_pimd_t_x *pimd) {
  if (!md_pimd_sync_host_x(pimd, 1)) return -1;
  return print_tri(out, "ENk", pimd->ENk, pimd->N);
}

int md_pimd_fprint_ENk2_x(FILE *out, md_pimd_t_x *pimd) {
  if (!md_pimd_sync_host_x(pimd, 1)) return -1;
  return print_tri(out, "ENk2", pimd->ENk2, pimd->N);
}

int md_pimd_fprint_VBN_x(FILE *out, md_pimd_t_x *pimd) {
  if (!md_pimd_sync_host_x(pimd, 1)) return -1;
  fprintf(out, "VBN %d\n", pimd->N);
  for (int i = 0; i <= pimd->N; ++i) fprintf(out, "%f ", pimd->VBN[i]);
  fprintf(out, "\n");
  return ferror(out);
}

}  // extern "C"

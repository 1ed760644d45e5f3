/*
 * Golden-vector generator.  Links the reference's own nocl library
 * (oracle/_ref/libxmd_ref.so, compiled from the reference sources by
 * oracle/Makefile) and writes inputs and outputs of every per-step stage for
 * small seeded systems into a binary fixture file:
 *
 *   record := char name[64] (NUL padded) | int64 count | double data[count]
 *
 * Run in the build container only (the reference tree is not shipped):
 *   make -C oracle ref && oracle/_ref/make_golden tests/golden/golden.bin
 *   oracle/_ref/make_golden tests/golden/configs.bin configs
 *   oracle/_ref/make_golden tests/golden/metric.bin metric
 */
#include "md_pimd_x.h"
#include "md_simulation_x.h"
#include "md_statistics_x.h"
#include "md_util_x.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static FILE *out;

static void dump(const char *name, const double *d, int64_t n) {
  char buf[64];
  memset(buf, 0, sizeof buf);
  strncpy(buf, name, 63);
  fwrite(buf, 1, 64, out);
  fwrite(&n, sizeof n, 1, out);
  fwrite(d, sizeof(double), (size_t)n, out);
}

static void dump1(const char *name, double v) { dump(name, &v, 1); }

static void dumpf(const char *prefix, const char *what, const double *d, int64_t n) {
  char name[64];
  snprintf(name, sizeof name, "%s/%s", prefix, what);
  dump(name, d, n);
}

/* particles -> flat arrays in reference bead order */
static void dump_beads(const char *prefix, const char *tag, md_simulation_t_x *sim) {
  int nb = sim->N;
  double *a = malloc(sizeof(double) * 3 * nb);
  char w[48];
  const char *fld[3] = {"x", "v", "f"};
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < nb; ++i)
      for (int d = 0; d < 3; ++d) {
        const md_particle_t_x *p = &sim->particles[i];
        a[3 * i + d] = k == 0 ? p->x[d] : (k == 1 ? p->v[d] : p->f[d]);
      }
    snprintf(w, sizeof w, "%s%s", fld[k], tag);
    dumpf(prefix, w, a, 3 * nb);
  }
  for (int i = 0; i < nb; ++i) a[i] = sim->particles[i].m;
  snprintf(w, sizeof w, "m%s", tag);
  dumpf(prefix, w, a, nb);
  free(a);
}

static void dump_forces(const char *prefix, const char *what, md_simulation_t_x *sim) {
  int nb = sim->N;
  double *a = malloc(sizeof(double) * 3 * nb);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) a[3 * i + d] = sim->particles[i].f[d];
  dumpf(prefix, what, a, 3 * nb);
  free(a);
}

static void dump_nhc(const char *prefix, const char *what, md_simulation_t_x *sim) {
  int rec = 1 + 3 * MD_NHC_LENGTH_X;
  double *a = malloc(sizeof(double) * rec * sim->Nf);
  for (int i = 0; i < sim->Nf; ++i) {
    double *r = &a[rec * i];
    r[0] = sim->nhcs[i].f;
    for (int j = 0; j < MD_NHC_LENGTH_X; ++j) {
      r[1 + j] = sim->nhcs[i].theta[j];
      r[1 + MD_NHC_LENGTH_X + j] = sim->nhcs[i].vtheta[j];
      r[1 + 2 * MD_NHC_LENGTH_X + j] = sim->nhcs[i].Q[j];
    }
  }
  dumpf(prefix, what, a, (int64_t)rec * sim->Nf);
  free(a);
}

static void dump_tri(const char *prefix, const char *what, md_num_t_x **T, int N) {
  int64_t n = (int64_t)N * (N + 1) / 2;
  dumpf(prefix, what, T[0], n);
}

static double stat_of(md_stats_t_x *s) { return s->es[s->N - 1]; }

typedef struct {
  const char *name;
  int N, P;
  double T, vi;
  int kind;          /* 0 trap+gaussian, 1 UEG ewald */
  int image;
  double L;          /* box edge (UEG) */
  int nsum;          /* Ewald cube half-width */
  int fgroup;        /* DOFs per thermostat chain */
  int steps;
  double dt;
  unsigned seed;
} gcase;

/* kind 0: trap + Gaussian; 1: UEG, Ewald; 2: periodic box, periodic Gaussian */
static md_pair_force_t_x pair_force_of(const gcase *c) {
  return c->kind == 0 ? md_gaussian_force_x : (c->kind == 1 ? md_coulomb_3d_ewald_force_x : md_periodic_gaussian_force_x);
}

static md_pair_energy_t_x pair_energy_of(const gcase *c) {
  return c->kind == 0 ? md_gaussian_energy_x
                      : (c->kind == 1 ? md_coulomb_3d_ewald_energy_x : md_periodic_gaussian_energy_x);
}

static void set_params(const gcase *c) {
  md_num_t_x zero[3] = {0, 0, 0};
  md_set_harmonic_trap_frequency_x(1.0);
  md_set_harmonic_trap_center_x(zero);
  md_set_gaussian_force_strength_x(3.0);
  md_set_guassian_force_range_x(0.5);
  md_set_coulomb_force_strength_x(1.0);
  md_set_coulomb_truc_x(5.0 / c->L);
  md_set_coulomb_n_sum_x(c->nsum);
}

static void full_force(const gcase *c, md_pimd_t_x *pimd) {
  md_reset_force_x(pimd->sim);
  if (c->kind == 0) md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
  md_pimd_calc_pair_force_x(pimd, pair_force_of(c));
  md_pimd_fast_fill_ENk_x(pimd, c->image);
  md_pimd_fill_VB_x(pimd);
  md_pimd_fast_fill_force_VB_x(pimd, c->image);
}

static void run_case(const gcase *c) {
  const char *pf = c->name;
  int N = c->N, P = c->P, nb = N * P;
  set_params(c);
  md_set_seed_x(c->seed);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, c->fgroup, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {c->L, c->L, c->L};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, c->vi, c->T);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  if (c->kind == 0) {
    /* positions i.i.d. Normal(0,1): particle l, bead j, then dimension */
    for (int l = 1; l <= N; ++l)
      for (int j = 1; j <= P; ++j)
        for (int d = 0; d < 3; ++d) sim->particles[MD_PIMD_INDEX_X(l, j, P)].x[d] = md_random_gaussian_x();
  } else {
    md_num_t_x ctr[3] = {c->L / 2, c->L / 2, c->L / 2};
    md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.4);
    md_periodic_boundary_x(sim);
  }
  md_init_maxwell_vel_x(sim, c->T);
  md_init_nhc_x(sim, md_get_quick_Q_x(c->T, pimd->omegaP), 0, sim->Nf, 0);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) sim->particles[i].f[d] = 0;

  double meta[10] = {N, P, c->T, c->vi, c->kind, c->image, c->L, c->nsum, c->fgroup, c->dt};
  dumpf(pf, "meta", meta, 10);
  dumpf(pf, "beta_omegaP", (double[]){pimd->beta, pimd->omegaP}, 2);
  dump_beads(pf, "0", sim);
  dump_nhc(pf, "nhc0", sim);

  /* stage forces, each from zero */
  if (c->kind == 0) {
    md_reset_force_x(sim);
    md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
    dump_forces(pf, "f_trap", sim);
  }
  md_reset_force_x(sim);
  md_pimd_calc_pair_force_x(pimd, pair_force_of(c));
  dump_forces(pf, "f_pair", sim);
  if (c->kind == 1) {
    md_reset_force_x(sim);
    md_pimd_calc_pair_force_x(pimd, md_coulomb_3d_ewald_force_R_x);
    dump_forces(pf, "f_pair_R", sim);
    md_reset_force_x(sim);
    md_pimd_calc_pair_force_x(pimd, md_coulomb_3d_ewald_force_NI_x);
    dump_forces(pf, "f_pair_NI", sim);
    dumpf(pf, "madelung", (double[]){md_calc_3d_ueg_madelung_x(L3)}, 1);
  }
  md_pimd_fill_ENk_x(pimd, c->image);
  dump_tri(pf, "ENk_slow", pimd->ENk, N);
  md_pimd_fast_fill_ENk_x(pimd, c->image);
  dump_tri(pf, "ENk", pimd->ENk, N);
  dumpf(pf, "Eint", pimd->Eint, N);
  md_pimd_fill_VB_x(pimd);
  dumpf(pf, "VBN", pimd->VBN, N + 1);
  md_reset_force_x(sim);
  md_pimd_fast_fill_force_VB_x(pimd, c->image);
  dump_forces(pf, "f_exch", sim);
  {
    /* single-value helpers: cycle energies E_n^(k) and link-virial terms for
     * (n, k) in {(N, 1), (N, N), (N-1, 2), (1, 1)}, the shifts xminE(n) for
     * every n, fast_xminE(u) for every u (V = VBN), xminE2 at xi' = 0.5 and 0,
     * and spring gradients of beads (l, j) on cycle (N, k = N). */
    int nk[4][2] = {{N, 1}, {N, N}, {N - 1, 2}, {1, 1}};
    double h[8];
    for (int q = 0; q < 4; ++q) {
      h[2 * q] = md_pimd_ENk_x(pimd, nk[q][0], nk[q][1], c->image);
      h[2 * q + 1] = md_pimd_ENk2_x(pimd, nk[q][0], nk[q][1], c->image);
    }
    dumpf(pf, "h_enk", h, 8);
    double *xm = malloc(sizeof(double) * 4 * N);
    for (int n = 1; n <= N; ++n) {
      xm[n - 1] = md_pimd_xminE_x(pimd, n);
      xm[N + n - 1] = md_pimd_fast_xminE_x(pimd, n, pimd->VBN);
      xm[2 * N + n - 1] = md_pimd_xminE2_x(pimd, n, pimd->VBN, 0.5);
      xm[3 * N + n - 1] = md_pimd_xminE2_x(pimd, n, pimd->VBN, 0.0);
    }
    dumpf(pf, "h_xmin", xm, 4 * N);
    free(xm);
    double gr[12];
    int lj[2][2] = {{1, 1}, {N, P}};
    for (int q = 0; q < 2; ++q) md_pimd_dENk_x(pimd, N, N, lj[q][0], lj[q][1], c->image, gr + 3 * q);
    /* explicit triples: (bead (1,2); (1,3), (1,1)) and (bead (N,P); bead (1,1), (N,P-1)) */
    md_pimd_fast_dENk_x(pimd, MD_PIMD_INDEX_X(1, 2, P), MD_PIMD_INDEX_X(1, 3 <= P ? 3 : 1, P),
                        MD_PIMD_INDEX_X(1, 1, P), c->image, gr + 6);
    md_pimd_fast_dENk_x(pimd, MD_PIMD_INDEX_X(N, P, P), MD_PIMD_INDEX_X(1, 1, P), MD_PIMD_INDEX_X(N, P - 1, P),
                        c->image, gr + 9);
    dumpf(pf, "h_grad", gr, 12);
    double v2[2] = {md_pimd_fast_ENk2_x(pimd, MD_PIMD_INDEX_X(1, 1, P), MD_PIMD_INDEX_X(N, P, P),
                                        MD_PIMD_INDEX_X(1, 1, P), c->image),
                    md_pimd_fast_ENk2_x(pimd, MD_PIMD_INDEX_X(2, 1, P), MD_PIMD_INDEX_X(1, P, P),
                                        MD_PIMD_INDEX_X(1, 1, P), c->image)};
    dumpf(pf, "h_enk2", v2, 2);
  }
  if (N <= 16) {
    md_reset_force_x(sim);
    md_pimd_fill_force_VB_x(pimd, c->image);
    dump_forces(pf, "f_exch_slow", sim);
  }

  /* estimators on the full force */
  full_force(c, pimd);
  dump_forces(pf, "f_total", sim);
  md_stats_t_x *st = md_alloc_stats_x(1);
  md_pimd_calc_VBN_energy_x(pimd, st);
  dumpf(pf, "e_VBN", (double[]){stat_of(st)}, 1);
  md_pimd_fast_fill_ENk2_x(pimd, c->image);
  dump_tri(pf, "ENk2", pimd->ENk2, N);
  st->es[0] = 0;
  md_pimd_calc_vi_virial_energy_x(pimd, st, c->image);
  dumpf(pf, "e_vir", (double[]){stat_of(st)}, 1);
  st->es[0] = 0;
  md_pimd_calc_vi_virial_pressure_x(pimd, st, c->image);
  dumpf(pf, "e_press", (double[]){stat_of(st)}, 1);
  md_pimd_fill_ENk2_x(pimd, c->image);
  dump_tri(pf, "ENk2_slow", pimd->ENk2, N);
  st->es[0] = 0;
  md_pimd_calc_vi_virial_energy_x(pimd, st, c->image);
  dumpf(pf, "e_vir_slow", (double[]){stat_of(st)}, 1);
  if (c->kind == 0) {
    st->es[0] = 0;
    md_pimd_calc_trap_energy_x(pimd, md_harmonic_trap_energy_x, st);
    dumpf(pf, "e_trap", (double[]){stat_of(st)}, 1);
  }
  st->es[0] = 0;
  md_pimd_calc_pair_energy_x(pimd, pair_energy_of(c), st);
  dumpf(pf, "e_pair", (double[]){stat_of(st)}, 1);
  if (c->kind == 1) {
    /* the Ewald energy splits R, R2 and NI */
    md_pair_energy_t_x sp[3] = {md_coulomb_3d_ewald_energy_R_x, md_coulomb_3d_ewald_energy_R2_x,
                                md_coulomb_3d_ewald_energy_NI_x};
    double es[3];
    for (int q = 0; q < 3; ++q) {
      st->es[0] = 0;
      md_pimd_calc_pair_energy_x(pimd, sp[q], st);
      es[q] = stat_of(st);
    }
    dumpf(pf, "e_pair_R_R2_NI", es, 3);
  }
  st->Ts[0] = 0;
  md_calc_temperature_x(sim, st);
  dumpf(pf, "e_temperature", (double[]){st->Ts[0]}, 1);
  st->es[0] = 0;
  md_calc_kinetic_x(sim, st);
  dumpf(pf, "e_kin", (double[]){stat_of(st)}, 1);
  st->es[0] = 0;
  md_calc_nhc_kinetic_x(sim, st);
  dumpf(pf, "e_nhc", (double[]){stat_of(st)}, 1);
  st->es[0] = 0;
  md_pimd_calc_virial_energy_x(pimd, st);
  dumpf(pf, "e_virial_plain", (double[]){stat_of(st)}, 1);
  st->es[0] = 0;
  md_pimd_calc_centroid_virial_energy_x(pimd, st, 3);
  dumpf(pf, "e_virial_centroid", (double[]){stat_of(st)}, 1);
  {
    double vis[4] = {-1.0, -0.5, 0.0, 0.5}, sgn[4];
    md_stats_t_x *es = md_alloc_stats_x(1), *ss = md_alloc_stats_x(1);
    for (int q = 0; q < 4; ++q) {
      es->es[0] = 1;
      ss->es[0] = 1;
      md_pimd_calc_vi_sign_x(pimd, es, ss, vis[q]);
      sgn[q] = ss->es[0];
    }
    dumpf(pf, "vi_sign_args", vis, 4);
    dumpf(pf, "vi_sign", sgn, 4);
  }

  /* trajectory of full MD steps (velocity Verlet with NHC) */
  double *ev = malloc(sizeof(double) * (c->steps > 0 ? c->steps : 1));
  for (int s = 0; s < c->steps; ++s) {
    md_update_nhc_VV3_1_x(sim, c->dt);
    if (c->kind != 0) md_periodic_boundary_x(sim);
    full_force(c, pimd);
    md_update_nhc_VV3_2_x(sim, c->dt);
    md_pimd_fast_fill_ENk2_x(pimd, c->image);
    st->es[0] = 0;
    md_pimd_calc_vi_virial_energy_x(pimd, st, c->image);
    ev[s] = stat_of(st);
  }
  if (c->steps > 0) {
    dump_beads(pf, "_end", sim);
    dump_nhc(pf, "nhc_end", sim);
    dumpf(pf, "e_vir_series", ev, c->steps);
  }
  free(ev);
}

static void rng_case(void) {
  md_set_seed_x(1234);
  double u[8], g[9];
  for (int i = 0; i < 8; ++i) u[i] = md_random_uniform_x(-1.0, 2.0);
  for (int i = 0; i < 9; ++i) g[i] = md_random_gaussian_x();
  dump("rng/uniform", u, 8);
  dump("rng/gaussian", g, 9);
  /* Box-Muller cache is now primed with one value: consume it. */
  md_random_gaussian_x();
  md_set_seed_x(77);
  md_simulation_t_x *sim = md_alloc_simulation_x(12, 3, 0, MD_ND_RECT_BOX_X);
  md_init_particle_mass_x(sim, 1.0, 0, 6);
  md_init_particle_mass_x(sim, 2.5, 6, 12);
  md_init_maxwell_vel_x(sim, 0.7);
  double v[36];
  for (int i = 0; i < 12; ++i)
    for (int d = 0; d < 3; ++d) v[3 * i + d] = sim->particles[i].v[d];
  dump("rng/maxwell_v", v, 36);
  md_set_seed_x(5);
  md_simulation_t_x *s2 = md_alloc_simulation_x(15, 1, 0, MD_ND_RECT_BOX_X);
  md_pimd_t_x *p2 = md_alloc_pimd_x(s2, 5, 3, 1.0, 1.0);
  md_num_t_x ctr[3] = {0.5, 1.0, -0.5}, L[3] = {2, 3, 4};
  md_pimd_init_particle_uniform_pos_x(p2, ctr, L, 0.1);
  double x[45];
  for (int i = 0; i < 15; ++i)
    for (int d = 0; d < 3; ++d) x[3 * i + d] = s2->particles[i].x[d];
  dump("rng/uniform_pos", x, 45);
}

/* Two species with the same P (md_pimd_calc_pair_force_2_x / _energy_2_x):
 * A = 5 particles (mass 1), B = 3 particles (mass 2), P = 4, positions
 * Normal(0,1) in particle, bead, dimension order (A then B); Gaussian and
 * periodic Gaussian (box 3 x 3.5 x 4). */
static void two_species_case(void) {
  const int NA = 5, NB = 3, P = 4;
  md_num_t_x L3[3] = {3.0, 3.5, 4.0}, zero[3] = {0, 0, 0};
  md_set_gaussian_force_strength_x(3.0);
  md_set_guassian_force_range_x(0.5);
  md_set_harmonic_trap_center_x(zero);
  md_set_seed_x(21);
  md_simulation_t_x *sa = md_alloc_simulation_x(NA * P, 1, 0, MD_ND_RECT_BOX_X);
  md_simulation_t_x *sb = md_alloc_simulation_x(NB * P, 1, 0, MD_ND_RECT_BOX_X);
  md_init_nd_rect_box_x(sa->box, L3);
  md_init_nd_rect_box_x(sb->box, L3);
  md_pimd_t_x *pa = md_alloc_pimd_x(sa, NA, P, 1.0, 1.0);
  md_pimd_t_x *pb = md_alloc_pimd_x(sb, NB, P, 1.0, 1.0);
  md_init_particle_mass_x(sa, 1.0, 0, NA * P);
  md_init_particle_mass_x(sb, 2.0, 0, NB * P);
  for (int i = 0; i < NA * P; ++i)
    for (int d = 0; d < 3; ++d) sa->particles[i].x[d] = md_random_gaussian_x();
  for (int i = 0; i < NB * P; ++i)
    for (int d = 0; d < 3; ++d) sb->particles[i].x[d] = md_random_gaussian_x();
  dump_beads("two", "A0", sa);
  dump_beads("two", "B0", sb);
  md_pair_force_t_x pfs[2] = {md_gaussian_force_x, md_periodic_gaussian_force_x};
  md_pair_energy_t_x pes[2] = {md_gaussian_energy_x, md_periodic_gaussian_energy_x};
  const char *tag[2] = {"g", "pg"};
  for (int k = 0; k < 2; ++k) {
    char w[32];
    md_reset_force_x(sa);
    md_reset_force_x(sb);
    md_pimd_calc_pair_force_2_x(pa, pb, pfs[k]);
    snprintf(w, sizeof w, "fA_%s", tag[k]);
    dump_forces("two", w, sa);
    snprintf(w, sizeof w, "fB_%s", tag[k]);
    dump_forces("two", w, sb);
    md_stats_t_x *st = md_alloc_stats_x(1);
    st->es[0] = 0;
    md_pimd_calc_pair_energy_2_x(pa, pb, pes[k], st);
    snprintf(w, sizeof w, "e_%s", tag[k]);
    dumpf("two", w, (double[]){stat_of(st)}, 1);
  }
}

/* Observables (density distribution, pair correlation, S(q)) of two seeded
 * configurations: periodic (N = 20, P = 6, box 5 x 5.5 x 6, uniform
 * positions) and open (Gaussian positions). */
static void observables_case(void) {
  const int N = 20, P = 6, nb = N * P;
  md_num_t_x L3[3] = {5.0, 5.5, 6.0}, ctr[3] = {2.5, 2.75, 3.0}, c0[3] = {0.1, -0.2, 0.3};
  md_set_seed_x(31);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, 1.0, 1.0);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  for (int pass = 0; pass < 2; ++pass) {
    const char *pf = pass == 0 ? "obs_pbc" : "obs_open";
    if (pass == 0) {
      md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.3);
      md_periodic_boundary_x(sim);
    } else {
      for (int i = 0; i < nb; ++i)
        for (int d = 0; d < 3; ++d) sim->particles[i].x[d] = md_random_gaussian_x();
    }
    dump_beads(pf, "0", sim);
    md_stats_t_x *a = md_alloc_stats_x(24), *b = md_alloc_stats_x(30), *c = md_alloc_stats_x(16);
    md_pimd_calc_density_distribution_x(pimd, a, 3.0, 1 - pass, pass == 0 ? ctr : c0);
    md_pimd_calc_pair_correlation_x(pimd, b, 3.5, 7.0, 4, 1 - pass);
    md_pimd_calc_Sk_structure_x(pimd, c, 0.2, 0.35, 3);
    dumpf(pf, "density", a->fxs[0], 24);
    dumpf(pf, "paircorr", b->fxs[0], 30);
    dumpf(pf, "sk", c->fxs[0], 16);
  }
}

/* Lennard-Jones and helium pair forces and energies, one species and two:
 * 27 particles on a jittered 3 x 3 x 3 lattice (spacing a, jitter 0.1 a per
 * bead coordinate, uniform from the seed), P = 4, cubic box 3 a.  Two-species
 * records split the same beads into A = particles 1..15 (mass 1) and
 * B = 16..27 (mass 2). */
static void lattice_pos(md_simulation_t_x *sim, int N, int P, int first, int count, double a) {
  for (int l = 1; l <= count; ++l) {
    int site = first + l - 1;
    int ix = site % 3, iy = (site / 3) % 3, iz = site / 9;
    for (int j = 1; j <= P; ++j) {
      md_particle_t_x *q = &sim->particles[MD_PIMD_INDEX_X(l, j, P)];
      q->x[0] = (ix + 0.5) * a + md_random_uniform_x(-0.1 * a, 0.1 * a);
      q->x[1] = (iy + 0.5) * a + md_random_uniform_x(-0.1 * a, 0.1 * a);
      q->x[2] = (iz + 0.5) * a + md_random_uniform_x(-0.1 * a, 0.1 * a);
    }
  }
  (void)N;
}

static void lj_he_case(void) {
  const int N = 27, P = 4, NA = 15, NB = 12;
  md_set_helium_parameters_x(10.8, 544850.4, 13.353384, 1.3732412, 0.4253785, 0.178100, 1.241314, 2.9673);
  const double spacing[2] = {1.25, 3.1};
  md_pair_force_t_x pfs[2] = {md_LJ_periodic_force_x, md_periodic_helium_force_x};
  md_pair_energy_t_x pes[2] = {md_LJ_periodic_energy_x, md_periodic_helium_energy_x};
  const char *pfx[2] = {"lj", "he"};
  for (int k = 0; k < 2; ++k) {
    const double a = spacing[k];
    md_num_t_x L3[3] = {3 * a, 3 * a, 3 * a};
    md_set_seed_x(41 + k);
    md_simulation_t_x *sim = md_alloc_simulation_x(N * P, 1, 0, MD_ND_RECT_BOX_X);
    md_init_nd_rect_box_x(sim->box, L3);
    md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, 1.0, 1.0);
    md_init_particle_mass_x(sim, 1.0, 0, N * P);
    lattice_pos(sim, N, P, 0, N, a);
    dump_beads(pfx[k], "0", sim);
    dumpf(pfx[k], "L", L3, 3);
    md_reset_force_x(sim);
    md_pimd_calc_pair_force_x(pimd, pfs[k]);
    dump_forces(pfx[k], "f_pair", sim);
    md_stats_t_x *st = md_alloc_stats_x(1);
    st->es[0] = 0;
    md_pimd_calc_pair_energy_x(pimd, pes[k], st);
    dumpf(pfx[k], "e_pair", (double[]){stat_of(st)}, 1);
    /* two species from the same sites */
    md_simulation_t_x *sa = md_alloc_simulation_x(NA * P, 1, 0, MD_ND_RECT_BOX_X);
    md_simulation_t_x *sb = md_alloc_simulation_x(NB * P, 1, 0, MD_ND_RECT_BOX_X);
    md_init_nd_rect_box_x(sa->box, L3);
    md_init_nd_rect_box_x(sb->box, L3);
    md_pimd_t_x *pa = md_alloc_pimd_x(sa, NA, P, 1.0, 1.0);
    md_pimd_t_x *pb = md_alloc_pimd_x(sb, NB, P, 1.0, 1.0);
    md_init_particle_mass_x(sa, 1.0, 0, NA * P);
    md_init_particle_mass_x(sb, 2.0, 0, NB * P);
    lattice_pos(sa, NA, P, 0, NA, a);
    lattice_pos(sb, NB, P, NA, NB, a);
    dump_beads(pfx[k], "A0", sa);
    dump_beads(pfx[k], "B0", sb);
    md_reset_force_x(sa);
    md_reset_force_x(sb);
    md_pimd_calc_pair_force_2_x(pa, pb, pfs[k]);
    dump_forces(pfx[k], "fA", sa);
    dump_forces(pfx[k], "fB", sb);
    st->es[0] = 0;
    md_pimd_calc_pair_energy_2_x(pa, pb, pes[k], st);
    dumpf(pfx[k], "e2", (double[]){stat_of(st)}, 1);
  }
}

/* The remaining observables on three seeded systems in one box (A: 7
 * particles, B: 5, C: 4; P = 6; uniform positions in 5 x 5.5 x 6). */
static void observables2_case(void) {
  const int NA = 7, NB = 5, NC = 4, P = 6;
  md_num_t_x L3[3] = {5.0, 5.5, 6.0}, ctr[3] = {2.5, 2.75, 3.0};
  md_set_seed_x(57);
  md_simulation_t_x *s[3];
  md_pimd_t_x *pm[3];
  const int Ns[3] = {NA, NB, NC};
  const char *tag[3] = {"A0", "B0", "C0"};
  for (int k = 0; k < 3; ++k) {
    s[k] = md_alloc_simulation_x(Ns[k] * P, 1, 0, MD_ND_RECT_BOX_X);
    md_init_nd_rect_box_x(s[k]->box, L3);
    pm[k] = md_alloc_pimd_x(s[k], Ns[k], P, 1.0, 1.0);
    md_init_particle_mass_x(s[k], 1.0, 0, Ns[k] * P);
    md_pimd_init_particle_uniform_pos_x(pm[k], ctr, L3, 0.5);
    md_periodic_boundary_x(s[k]);
    dump_beads("obs2", tag[k], s[k]);
  }
  md_stats_t_x *h;
  h = md_alloc_stats_x(36);
  md_pimd_calc_ITCF_x(pm[0], h, 0.5, 4.5, 3);
  dumpf("obs2", "itcf", h->fxs[0], 36);
  h = md_alloc_stats_x(36);
  md_pimd_calc_ITCF_2_x(pm[0], pm[1], h, 0.5, 4.5, 3);
  dumpf("obs2", "itcf2", h->fxs[0], 36);
  h = md_alloc_stats_x(20);
  md_pimd_calc_pair_correlation_2_x(pm[0], pm[1], h, 3.5, 7.0, 2, 1);
  dumpf("obs2", "paircorr2", h->fxs[0], 20);
  h = md_alloc_stats_x(20);
  md_pimd_calc_pair_correlation_2_x(pm[0], pm[1], h, 6.0, 3.0, 5, 0);
  dumpf("obs2", "paircorr2_open", h->fxs[0], 20);
  {
    md_num_t_x box[4] = {0.5, 4.5, 0.5, 5.0};
    h = md_alloc_stats_x(20);
    md_pimd_calc_pair_correlation_2d_x(pm[0], h, 3.0, 5.0, box);
    dumpf("obs2", "paircorr_2d", h->fxs[0], 20);
  }
  h = md_alloc_stats_x(25);
  md_pimd_calc_2d_pair_correlation_x(pm[0], h, 3.0, 3.0, 4, 1);
  dumpf("obs2", "pc2d", h->fxs[0], 25);
  h = md_alloc_stats_x(25);
  md_pimd_calc_2d_pair_correlation_2_x(pm[0], pm[1], h, 3.0, 3.0, 4, 0);
  dumpf("obs2", "pc2d_2", h->fxs[0], 25);
  h = md_alloc_stats_x(16);
  md_pimd_calc_Sk_structure_2_x(pm[0], pm[1], h, 0.2, 0.35, 3);
  dumpf("obs2", "sk2", h->fxs[0], 16);
  h = md_alloc_stats_x(16);
  md_pimd_calc_Sk_structure_full_2_x(pm[0], pm[1], h, 0.2, 0.35, 3);
  dumpf("obs2", "sk_full2", h->fxs[0], 16);
  md_sk_sphere_t_x *sk = md_alloc_sk_sphere_3d_x(0.0, 0.6, 2.0);
  dumpf("obs2", "sphere_qs", sk->qs, sk->N);
  dumpf("obs2", "sphere_qis", sk->qis, 3 * sk->N);
  h = md_alloc_stats_x(sk->N + 3);
  md_pimd_calc_sphere_Sk_structure_x(pm[0], h, 3, sk);
  dumpf("obs2", "sphere_sk", h->fxs[0], sk->N + 3);
  h = md_alloc_stats_x(sk->N);
  md_pimd_calc_sphere_Sk_structure_full_2_x(pm[0], pm[1], h, 8, sk);
  dumpf("obs2", "sphere_sk_full2", h->fxs[0], sk->N);
  h = md_alloc_stats_x(sk->N - 2);
  md_pimd_calc_sphere_Sk_structure_3_x(pm[0], pm[1], pm[2], h, 2, sk);
  dumpf("obs2", "sphere_sk3", h->fxs[0], sk->N - 2);
}

/* Full-size parity cases at the benchmark configurations: trapped bosons
 * N=256, P=64, beta=2 (g=3, s=0.5) at xi in {0, 0.25, 0.5, 0.75, 1}, and the
 * spin-polarised UEG N=128, P=64, r_s=2, theta=0.5, alpha=5/L, n=8 at xi in
 * {0, 0.5, 1}.  One seeded initial state per system (x, v stored whole); per xi the
 * total force at that state (every `stride`-th bead) with the fast and with
 * the cycle-walk exchange algorithms, V_B^(N) and the
 * thermodynamic and virial estimators; for the last xi `steps` MD steps
 * (dt, NHC M=8) and the final positions and velocities (every stride-th bead). */
static void dump_sub(const char *prefix, const char *what, md_simulation_t_x *sim, int field, int stride) {
  int nb = sim->N, n = 0;
  double *a = malloc(sizeof(double) * 3 * nb);
  for (int i = 0; i < nb; i += stride)
    for (int d = 0; d < 3; ++d) {
      const md_particle_t_x *q = &sim->particles[i];
      a[n++] = field == 0 ? q->x[d] : (field == 1 ? q->v[d] : q->f[d]);
    }
  dumpf(prefix, what, a, n);
  free(a);
}

static void config_case(const gcase *c, const double *xis, int nxi, int stride) {
  const char *pf = c->name;
  int N = c->N, P = c->P, nb = N * P;
  set_params(c);
  md_set_seed_x(c->seed);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {c->L, c->L, c->L};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, xis[0], c->T);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  if (c->kind == 0) {
    for (int l = 1; l <= N; ++l)
      for (int j = 1; j <= P; ++j)
        for (int d = 0; d < 3; ++d) sim->particles[MD_PIMD_INDEX_X(l, j, P)].x[d] = md_random_gaussian_x();
  } else {
    md_num_t_x ctr[3] = {c->L / 2, c->L / 2, c->L / 2};
    md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.4);
    md_periodic_boundary_x(sim);
  }
  md_init_maxwell_vel_x(sim, c->T);
  md_init_nhc_x(sim, md_get_quick_Q_x(c->T, pimd->omegaP), 0, sim->Nf, 0);
  double meta[10] = {N, P, c->T, 0, c->kind, c->image, c->L, c->nsum, stride, c->dt};
  dumpf(pf, "meta", meta, 10);
  dumpf(pf, "xis", xis, nxi);
  /* positions and velocities of every bead; the chains start from
   * md_init_nhc_x as above (deterministic) */
  dump_sub(pf, "x0", sim, 0, 1);
  dump_sub(pf, "v0", sim, 1, 1);
  /* the pair (and trap) force is xi-independent: once */
  md_reset_force_x(sim);
  if (c->kind == 0) md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
  md_pimd_calc_pair_force_x(pimd, pair_force_of(c));
  double *f0 = malloc(sizeof(double) * 3 * nb);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) f0[3 * i + d] = sim->particles[i].f[d];
  dump_sub(pf, "f_ext", sim, 2, stride);
  md_stats_t_x *st = md_alloc_stats_x(1);
  st->es[0] = 0;
  md_pimd_calc_pair_energy_x(pimd, pair_energy_of(c), st); /* xi-independent */
  const double e_pair = stat_of(st);
  for (int q = 0; q < nxi; ++q) {
    char w[32];
    pimd->vi = xis[q];
    /* the cycle-walk algorithm first (its disagreement with the fast one
     * measures the reference's own reproducibility on this input) */
    for (int i = 0; i < nb; ++i)
      for (int d = 0; d < 3; ++d) sim->particles[i].f[d] = f0[3 * i + d];
    md_pimd_fill_ENk_x(pimd, c->image);
    md_pimd_fill_VB_x(pimd);
    md_pimd_fill_force_VB_x(pimd, c->image);
    snprintf(w, sizeof w, "f_total_slow_%d", q);
    dump_sub(pf, w, sim, 2, stride);
    for (int i = 0; i < nb; ++i)
      for (int d = 0; d < 3; ++d) sim->particles[i].f[d] = f0[3 * i + d];
    md_pimd_fast_fill_ENk_x(pimd, c->image);
    md_pimd_fill_VB_x(pimd);
    md_pimd_fast_fill_force_VB_x(pimd, c->image);
    snprintf(w, sizeof w, "f_total_%d", q);
    dump_sub(pf, w, sim, 2, stride);
    double e[4];
    e[0] = pimd->VBN[N];
    st->es[0] = 0;
    md_pimd_calc_VBN_energy_x(pimd, st);
    e[1] = stat_of(st);
    md_pimd_fast_fill_ENk2_x(pimd, c->image);
    st->es[0] = 0;
    md_pimd_calc_vi_virial_energy_x(pimd, st, c->image);
    e[2] = stat_of(st);
    e[3] = e_pair;
    snprintf(w, sizeof w, "e_%d", q);
    dumpf(pf, w, e, 4);
  }
  /* trajectory at the last xi, starting from the total force above */
  double *ev = malloc(sizeof(double) * (c->steps > 0 ? c->steps : 1));
  for (int s = 0; s < c->steps; ++s) {
    md_update_nhc_VV3_1_x(sim, c->dt);
    if (c->kind != 0) md_periodic_boundary_x(sim);
    full_force(c, pimd);
    md_update_nhc_VV3_2_x(sim, c->dt);
    md_pimd_fast_fill_ENk2_x(pimd, c->image);
    st->es[0] = 0;
    md_pimd_calc_vi_virial_energy_x(pimd, st, c->image);
    ev[s] = stat_of(st);
  }
  if (c->steps > 0) {
    dump_sub(pf, "x_end", sim, 0, stride);
    dump_sub(pf, "v_end", sim, 1, stride);
    dumpf(pf, "e_vir_series", ev, c->steps);
  }
  free(ev);
  free(f0);
}

/* Headline configuration (the benchmark workload): spin-polarised UEG
 * N=1024, P=32, r_s=2, theta=0.5, xi=1, alpha=5/L, n=8, set up from seed 1
 * exactly as the benchmark sets it up.  Records:
 *   x0          every bead's position (the seeded state, for a bit-exact check)
 *   enk_rows    rows n of the cycle-energy table E_n^(1..n) for n in ENK_ROWS
 *   VBN         V_B^(0..N)
 *   fx_end      exchange + ring-spring force (md_reset_force_x, then the fast
 *               path md_pimd_fast_fill_ENk_x / _fill_VB_x /
 *               _fast_fill_force_VB_x) of beads 1 and P of every particle
 *   fx_mid      the same for every 31st bead (interior springs)
 *   e_est       thermodynamic (md_pimd_calc_VBN_energy_x) and virial
 *               (md_pimd_fast_fill_ENk2_x + md_pimd_calc_vi_virial_energy_x)
 *               estimators at that force state
 *   f_slice_j   full Ewald pair force of bead slice j (1-based) for j in
 *               SLICES: the reference's md_pimd_calc_pair_force_x on a P=1
 *               object holding that slice's positions (so F itself, not F/P)
 *   e_slice_j   md_pimd_calc_pair_energy_x of the same P=1 object
 * The four slice evaluations (N(N-1)/2 pairs of two 17^3 image cubes each,
 * about 150 s per evaluation on one core) run on four threads; the
 * reference's interaction callbacks only read the process globals. */
#include <pthread.h>

enum { M_N = 1024, M_P = 32, M_NSLICE = 2 };
static const int SLICES[M_NSLICE] = {1, 17};
static const int ENK_ROWS[] = {1, 2, 63, 64, 65, 129, 500, 1023, 1024};

typedef struct {
  md_simulation_t_x *src;
  int slice, energy;
  double L, T;
  double *out; /* 3N forces or 1 energy */
} slice_job;

static void *slice_eval(void *arg) {
  slice_job *J = arg;
  md_simulation_t_x *sim = md_alloc_simulation_x(M_N, 1, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {J->L, J->L, J->L};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, M_N, 1, 1.0, J->T);
  md_init_particle_mass_x(sim, 1.0, 0, M_N);
  for (int l = 1; l <= M_N; ++l)
    for (int d = 0; d < 3; ++d)
      sim->particles[l - 1].x[d] = J->src->particles[MD_PIMD_INDEX_X(l, J->slice, M_P)].x[d];
  if (J->energy) {
    md_stats_t_x *st = md_alloc_stats_x(1);
    st->es[0] = 0;
    md_pimd_calc_pair_energy_x(pimd, md_coulomb_3d_ewald_energy_x, st);
    J->out[0] = st->es[0];
  } else {
    md_reset_force_x(sim);
    md_pimd_calc_pair_force_x(pimd, md_coulomb_3d_ewald_force_x);
    for (int l = 0; l < M_N; ++l)
      for (int d = 0; d < 3; ++d) J->out[3 * l + d] = sim->particles[l].f[d];
  }
  return NULL;
}

static void metric_case(void) {
  const char *pf = "cfg_ueg1024";
  const int N = M_N, P = M_P, nb = N * P;
  const double rs = 2.0, L = pow(4 * M_PI * N / 3.0, 1.0 / 3.0) * rs;
  const double kF = pow(6 * M_PI * M_PI * N / (L * L * L), 1.0 / 3.0);
  /*           name N  P  T             vi   kind img L  nsum fgrp steps dt  seed */
  gcase c = {pf, N, P, 0.25 * kF * kF, 1.0, 1, 1, L, 8, 1, 0, 0.05, 1};
  set_params(&c);
  md_set_seed_x(c.seed);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {L, L, L};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, c.vi, c.T);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  md_num_t_x ctr[3] = {L / 2, L / 2, L / 2};
  md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.4);
  md_periodic_boundary_x(sim);
  md_init_maxwell_vel_x(sim, c.T);
  double meta[10] = {N, P, c.T, c.vi, c.kind, c.image, L, c.nsum, 1, c.dt};
  dumpf(pf, "meta", meta, 10);
  dump_sub(pf, "x0", sim, 0, 1);
  /* slice evaluations in the background while the exchange path runs */
  slice_job jobs[2 * M_NSLICE];
  pthread_t th[2 * M_NSLICE];
  for (int q = 0; q < 2 * M_NSLICE; ++q) {
    jobs[q].src = sim;
    jobs[q].slice = SLICES[q % M_NSLICE];
    jobs[q].energy = q >= M_NSLICE;
    jobs[q].L = L;
    jobs[q].T = c.T;
    jobs[q].out = malloc(sizeof(double) * 3 * N);
  }
  /* the slice objects copy positions before the exchange path changes f
   * only; positions stay untouched, so the threads may read them */
  for (int q = 0; q < 2 * M_NSLICE; ++q) pthread_create(&th[q], NULL, slice_eval, &jobs[q]);
  md_reset_force_x(sim);
  md_pimd_fast_fill_ENk_x(pimd, 1);
  md_pimd_fill_VB_x(pimd);
  md_pimd_fast_fill_force_VB_x(pimd, 1);
  int nrows = (int)(sizeof ENK_ROWS / sizeof ENK_ROWS[0]), tot = 0;
  for (int r = 0; r < nrows; ++r) tot += ENK_ROWS[r];
  double *rows = malloc(sizeof(double) * tot);
  int w = 0;
  for (int r = 0; r < nrows; ++r)
    for (int k = 1; k <= ENK_ROWS[r]; ++k) rows[w++] = pimd->ENk[ENK_ROWS[r] - 1][k - 1];
  double rowid[sizeof ENK_ROWS / sizeof ENK_ROWS[0]];
  for (int r = 0; r < nrows; ++r) rowid[r] = ENK_ROWS[r];
  dumpf(pf, "enk_row_ids", rowid, nrows);
  dumpf(pf, "enk_rows", rows, tot);
  dumpf(pf, "VBN", pimd->VBN, N + 1);
  double *fe = malloc(sizeof(double) * 6 * N);
  for (int e = 0; e < 2; ++e)
    for (int l = 1; l <= N; ++l)
      for (int d = 0; d < 3; ++d) fe[(e * N + l - 1) * 3 + d] = sim->particles[MD_PIMD_INDEX_X(l, e ? P : 1, P)].f[d];
  dumpf(pf, "fx_end", fe, 6 * N);
  dump_sub(pf, "fx_mid", sim, 2, 31);
  md_stats_t_x *st = md_alloc_stats_x(1);
  double e_est[2];
  st->es[0] = 0;
  md_pimd_calc_VBN_energy_x(pimd, st);
  e_est[0] = stat_of(st);
  md_pimd_fast_fill_ENk2_x(pimd, 1);
  st->es[0] = 0;
  md_pimd_calc_vi_virial_energy_x(pimd, st, 1);
  e_est[1] = stat_of(st);
  dumpf(pf, "e_est", e_est, 2);
  for (int q = 0; q < 2 * M_NSLICE; ++q) {
    pthread_join(th[q], NULL);
    char wname[32];
    snprintf(wname, sizeof wname, "%s_%d", jobs[q].energy ? "e_slice" : "f_slice", jobs[q].slice);
    dumpf(pf, wname, jobs[q].out, jobs[q].energy ? 1 : 3 * N);
    free(jobs[q].out);
  }
  free(rows);
  free(fe);
}

/* One replica of the batched-replica table, set up exactly as the engine's
 * test helper sets it up (seed, draw order, N(0,1) trap positions or UEG
 * uniform particle centres with beads scattered by +-0.4, Maxwell
 * velocities, chains), and its forces at that state: trap + pair
 * (f_ext) and trap + pair + fast-path exchange (f_total), every `stride`-th
 * bead, with the positions of those beads.  UEG pair forces use the
 * reference's Ewald at alpha = 5/L, n = 8 (converged, the same sum as any
 * other converged split such as the table's alpha = 2, |m| <= 200). */
static void replica_case(const char *name, int kind, int N, int P, double xi, double T, double L, unsigned seed,
                         int stride) {
  gcase c = {name, N, P, T, xi, kind, kind, L, 8, 1, 0, 0.01, seed};
  set_params(&c);
  md_init_config_x cfg = {1, 0, 0, seed, ""};
  md_init_setup_x(NULL, &cfg);
  md_set_seed_x(seed);
  const int nb = N * P;
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {L, L, L};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, xi, T);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  if (kind == 0) {
    for (int l = 1; l <= N; ++l)
      for (int j = 1; j <= P; ++j)
        for (int d = 0; d < 3; ++d) sim->particles[MD_PIMD_INDEX_X(l, j, P)].x[d] = md_random_gaussian_x();
  } else {
    md_num_t_x ctr[3] = {L / 2, L / 2, L / 2};
    md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.4);
    md_periodic_boundary_x(sim);
  }
  md_init_maxwell_vel_x(sim, T);
  double meta[10] = {N, P, T, xi, kind, kind, L, 8, stride, seed};
  dumpf(name, "meta", meta, 10);
  dump_sub(name, "x0", sim, 0, stride);
  md_reset_force_x(sim);
  if (kind == 0) md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
  md_pimd_calc_pair_force_x(pimd, pair_force_of(&c));
  dump_sub(name, "f_ext", sim, 2, stride);
  md_pimd_fast_fill_ENk_x(pimd, kind);
  md_pimd_fill_VB_x(pimd);
  md_pimd_fast_fill_force_VB_x(pimd, kind);
  dump_sub(name, "f_total", sim, 2, stride);
}

int main(int argc, char **argv) {
  const char *path = argc > 1 ? argv[1] : "golden.bin";
  out = fopen(path, "wb");
  if (!out) return 1;
  md_init_config_x cfg = {1, 0, 0, 1, ""};
  md_init_setup_x(NULL, &cfg);
  md_set_pimd_mode_x(1);
  if (argc > 2 && strcmp(argv[2], "replica") == 0 && argc == 12) {
    /* make_golden out.bin replica name kind N P xi T L seed stride */
    replica_case(argv[3], atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), atof(argv[7]), atof(argv[8]), atof(argv[9]),
                 (unsigned)strtoul(argv[10], NULL, 10), atoi(argv[11]));
    fclose(out);
    return 0;
  }
  if (argc > 2 && strcmp(argv[2], "metric") == 0) {
    metric_case();
    fclose(out);
    return 0;
  }
  if (argc > 2 && strcmp(argv[2], "configs") == 0) {
    double rs = 2.0, L128 = pow(4 * M_PI * 128 / 3.0, 1.0 / 3.0) * rs;
    double kF = pow(6 * M_PI * M_PI * 128 / (L128 * L128 * L128), 1.0 / 3.0);
    /*            name       N    P   T                  vi kind img L     nsum fgrp steps dt    seed */
    gcase trap = {"cfg_trap256", 256, 64, 0.5, 0, 0, 0, 10.0, 2, 1, 3, 0.01, 12};
    gcase ueg = {"cfg_ueg128", 128, 64, 0.25 * kF * kF, 0, 1, 1, L128, 8, 1, 2, 0.05, 13};
    const double xt[5] = {0.0, 0.25, 0.5, 0.75, 1.0}, xu[3] = {0.0, 0.5, 1.0};
    config_case(&trap, xt, 5, 13);
    config_case(&ueg, xu, 3, 7);
    fclose(out);
    return 0;
  }
  rng_case();
  double rs = 2.0;
  double L8 = pow(4 * M_PI * 8 / 3.0, 1.0 / 3.0) * rs;
  double L16 = pow(4 * M_PI * 16 / 3.0, 1.0 / 3.0) * rs;
  /* UEG temperature: theta = 0.5 of T_F = k_F^2 / 2, k_F = (6 pi^2 n)^(1/3) */
  double kF8 = pow(6 * M_PI * M_PI * 8 / (L8 * L8 * L8), 1.0 / 3.0);
  double kF16 = pow(6 * M_PI * M_PI * 16 / (L16 * L16 * L16), 1.0 / 3.0);
  gcase cases[] = {
      /* name         N   P   T      vi   kind img L    nsum fgrp steps dt    seed */
      {"trap16",      16, 32, 1.0,   1.0, 0, 0, 10.0, 2, 1, 20, 0.01, 1},
      {"trap16_g3",   16, 32, 1.0,   1.0, 0, 0, 10.0, 2, 3, 5,  0.01, 11},
      {"trap8_xi0",   8,  8,  0.5,   0.0, 0, 0, 10.0, 2, 1, 5,  0.01, 2},
      {"trap8_xi05",  8,  8,  0.5,   0.5, 0, 0, 10.0, 2, 1, 5,  0.01, 3},
      {"trap24_xi025",24, 6,  0.5,   0.25,0, 0, 10.0, 2, 7, 3,  0.01, 4},
      {"ueg8_n2",     8,  4,  0.25 * kF8 * kF8, 1.0, 1, 1, L8, 2, 1, 5, 0.05, 5},
      {"ueg8_n4_xi0", 8,  4,  0.25 * kF8 * kF8, 0.0, 1, 1, L8, 4, 1, 3, 0.05, 6},
      {"ueg16_n8",    16, 2,  0.25 * kF16 * kF16, 1.0, 1, 1, L16, 8, 1, 0, 0.05, 7},
  };
  for (size_t i = 0; i < sizeof cases / sizeof cases[0]; ++i) run_case(&cases[i]);
  two_species_case();
  observables_case();
  lj_he_case();
  observables2_case();
  {
    /* single-species periodic Gaussian (g=3, s=0.5) in a box of edge 3
     * with minimum-image springs, no trap */
    gcase pg = {"pgauss12", 12, 8, 1.0, 0.5, 2, 1, 3.0, 2, 1, 5, 0.01, 9};
    run_case(&pg);
  }
  fclose(out);
  return 0;
}

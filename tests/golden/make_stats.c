/*
 * Long-run estimator statistics of the reference's own nocl build, for the
 * statistical parity test (tests/test_gpu_stats.py).  Trapped bosons N = 16,
 * P = 32, omega = 1, beta = 1, xi = 1, Gaussian pair g = 3, s = 0.5,
 * dt = 0.01, massive NHC (the reference's chain length), seed 7; positions
 * i.i.d. Normal(0,1) per bead coordinate, Maxwell velocities.  After
 * `equil` steps, every step's estimators are averaged in `nblock` blocks:
 *   vir   md_pimd_calc_vi_virial_energy_x (after md_pimd_fast_fill_ENk2_x)
 *   vbn   md_pimd_calc_VBN_energy_x
 *   trap  md_pimd_calc_trap_energy_x
 *   pair  md_pimd_calc_pair_energy_x
 * Output (JSON): per estimator the mean, the standard error of the mean
 * from the block means, and the block means themselves (for coarser
 * blockings).
 *
 * Run in the build container only (links oracle/_ref/libxmd_ref.so):
 *   oracle/_ref/make_stats 10000 1000000 100 > tests/golden/stats_trap16.json
 */
#include "md_pimd_x.h"
#include "md_simulation_x.h"
#include "md_statistics_x.h"
#include "md_util_x.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static double take(md_stats_t_x *st) {
  double v = st->es[st->N - 1];
  st->es[st->N - 1] = 0;
  return v;
}

int main(int argc, char **argv) {
  const int equil = argc > 1 ? atoi(argv[1]) : 10000;
  const int steps = argc > 2 ? atoi(argv[2]) : 100000;
  const int nblock = argc > 3 ? atoi(argv[3]) : 100;
  const int N = 16, P = 32, nb = N * P;
  const double T = 1.0, xi = 1.0, dt = 0.01;
  md_num_t_x zero[3] = {0, 0, 0}, L3[3] = {10, 10, 10};
  md_set_harmonic_trap_frequency_x(1.0);
  md_set_harmonic_trap_center_x(zero);
  md_set_gaussian_force_strength_x(3.0);
  md_set_guassian_force_range_x(0.5);
  md_set_seed_x(7);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, xi, T);
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  for (int l = 1; l <= N; ++l)
    for (int j = 1; j <= P; ++j)
      for (int d = 0; d < 3; ++d) sim->particles[MD_PIMD_INDEX_X(l, j, P)].x[d] = md_random_gaussian_x();
  md_init_maxwell_vel_x(sim, T);
  md_init_nhc_x(sim, md_get_quick_Q_x(T, pimd->omegaP), 0, sim->Nf, 0);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) sim->particles[i].f[d] = 0;

  md_stats_t_x *st = md_alloc_stats_x(1);
  st->es[0] = 0;
  const char *names[4] = {"vir", "vbn", "trap", "pair"};
  double *blk = calloc((size_t)4 * nblock, sizeof(double));
  const int per = steps / nblock;
  md_reset_force_x(sim);
  md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
  md_pimd_calc_pair_force_x(pimd, md_gaussian_force_x);
  md_pimd_fast_fill_ENk_x(pimd, 0);
  md_pimd_fill_VB_x(pimd);
  md_pimd_fast_fill_force_VB_x(pimd, 0);
  for (int s = 0; s < equil + per * nblock; ++s) {
    md_update_nhc_VV3_1_x(sim, dt);
    md_reset_force_x(sim);
    md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
    md_pimd_calc_pair_force_x(pimd, md_gaussian_force_x);
    md_pimd_fast_fill_ENk_x(pimd, 0);
    md_pimd_fill_VB_x(pimd);
    md_pimd_fast_fill_force_VB_x(pimd, 0);
    md_update_nhc_VV3_2_x(sim, dt);
    if (s < equil) continue;
    const int b = (s - equil) / per;
    double e[4];
    md_pimd_fast_fill_ENk2_x(pimd, 0);
    md_pimd_calc_vi_virial_energy_x(pimd, st, 0);
    e[0] = take(st);
    md_pimd_calc_VBN_energy_x(pimd, st);
    e[1] = take(st);
    md_pimd_calc_trap_energy_x(pimd, md_harmonic_trap_energy_x, st);
    e[2] = take(st);
    md_pimd_calc_pair_energy_x(pimd, md_gaussian_energy_x, st);
    e[3] = take(st);
    for (int k = 0; k < 4; ++k) blk[k * nblock + b] += e[k] / per;
  }
  printf("{\n  \"system\": \"trapped bosons N=16 P=32 beta=1 xi=1 g=3 s=0.5 dt=0.01 seed=7\",\n");
  printf("  \"equil\": %d, \"steps\": %d, \"nblock\": %d, \"nhc_length\": %d,\n", equil, per * nblock, nblock,
         MD_NHC_LENGTH_X);
  for (int k = 0; k < 4; ++k) {
    double m = 0, v = 0;
    for (int b = 0; b < nblock; ++b) m += blk[k * nblock + b];
    m /= nblock;
    for (int b = 0; b < nblock; ++b) v += (blk[k * nblock + b] - m) * (blk[k * nblock + b] - m);
    v /= (nblock - 1);
    printf("  \"%s\": {\"mean\": %.17g, \"sem\": %.17g, \"blocks\": [", names[k], m, sqrt(v / nblock));
    for (int b = 0; b < nblock; ++b) printf("%.17g%s", blk[k * nblock + b], b < nblock - 1 ? ", " : "");
    printf("]}%s\n", k < 3 ? "," : "");
  }
  printf("}\n");
  free(blk);
  return 0;
}

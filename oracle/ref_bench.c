/*
 * ORACLE side -- CPU baseline only.  Times the reference's own nocl library
 * (oracle/_ref/libxmd_ref.so, single thread) on the per-step path of the
 * uniform-electron-gas workload, on a bounded sample:
 *
 *   - pair stage: md_coulomb_3d_ewald_force_x (the callback the reference's
 *     md_pimd_calc_pair_force_x loop calls, md_pimd_x.c:1660-1674) timed on
 *     the first `pairs` pairs of slice 1; the step needs N(N-1)/2 * P pairs,
 *     each with the same cube of (2n+1)^3 terms, so the per-pair time scales
 *     exactly;
 *   - exchange stage (md_pimd_fast_fill_ENk_x, md_pimd_fill_VB_x,
 *     md_pimd_fast_fill_force_VB_x) and the integrator halves + wrap + reset,
 *     timed once each at the full size.
 *
 * usage: ref_bench N P rs theta xi nsum pairs seed
 * prints one line: key=value pairs (seconds, pair counts, estimated pbs/s).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "md_pimd_x.h"
#include "md_simulation_x.h"
#include "md_statistics_x.h"
#include "md_util_x.h"

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv) {
  int N = argc > 1 ? atoi(argv[1]) : 1024;
  int P = argc > 2 ? atoi(argv[2]) : 32;
  double rs = argc > 3 ? atof(argv[3]) : 2.0;
  double theta = argc > 4 ? atof(argv[4]) : 0.5;
  double xi = argc > 5 ? atof(argv[5]) : 1.0;
  int nsum = argc > 6 ? atoi(argv[6]) : 8;
  long pairs = argc > 7 ? atol(argv[7]) : 2000;
  unsigned seed = argc > 8 ? (unsigned)atoi(argv[8]) : 1;

  double L = pow(4 * M_PI * N / 3.0, 1.0 / 3.0) * rs;
  double kF = pow(6 * M_PI * M_PI * N / (L * L * L), 1.0 / 3.0);
  double T = theta * kF * kF / 2;
  md_init_config_x cfg = {1, 0, 0, seed, ""};
  md_init_setup_x(NULL, &cfg);
  md_set_pimd_mode_x(1);
  md_set_coulomb_force_strength_x(1.0);
  md_set_coulomb_truc_x(5.0 / L);
  md_set_coulomb_n_sum_x(nsum);
  md_simulation_t_x *sim = md_alloc_simulation_x(N * P, 1, 0, MD_ND_RECT_BOX_X);
  md_num_t_x L3[3] = {L, L, L}, ctr[3] = {L / 2, L / 2, L / 2};
  md_init_nd_rect_box_x(sim->box, L3);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, xi, T);
  md_init_particle_mass_x(sim, 1.0, 0, N * P);
  md_pimd_init_particle_uniform_pos_x(pimd, ctr, L3, 0.4);
  md_periodic_boundary_x(sim);
  md_init_maxwell_vel_x(sim, T);
  md_init_nhc_x(sim, md_get_quick_Q_x(T, pimd->omegaP), 0, sim->Nf, 0);

  /* pair sample, slice 1 */
  long total_slice = (long)N * (N - 1) / 2, done = 0;
  double f[3] = {0, 0, 0};
  double t0 = now();
  for (int l = 2; l <= N && done < pairs; ++l)
    for (int l2 = 1; l2 < l && done < pairs; ++l2, ++done)
      md_coulomb_3d_ewald_force_x(sim->particles[MD_PIMD_INDEX_X(l, 1, P)].x,
                                  sim->particles[MD_PIMD_INDEX_X(l2, 1, P)].x, f, L3, 1.0);
  double t_pairs = now() - t0;
  double per_pair = t_pairs / (double)done;

  /* exchange stage, full size */
  t0 = now();
  md_pimd_fast_fill_ENk_x(pimd, 1);
  md_pimd_fill_VB_x(pimd);
  md_pimd_fast_fill_force_VB_x(pimd, 1);
  double t_exch = now() - t0;

  /* integrator + wrap + reset, full size */
  t0 = now();
  md_update_nhc_VV3_1_x(sim, 0.05);
  md_periodic_boundary_x(sim);
  md_reset_force_x(sim);
  md_update_nhc_VV3_2_x(sim, 0.05);
  double t_int = now() - t0;

  double t_step = per_pair * (double)total_slice * P + t_exch + t_int;
  printf("pairs_timed=%ld t_pairs=%.6f per_pair=%.9e t_exchange=%.6f t_integrator=%.6f "
         "t_step_est=%.3f pbs_per_s=%.6e fsum=%.3e\n",
         done, t_pairs, per_pair, t_exch, t_int, t_step, (double)N * P / t_step, f[0] + f[1] + f[2]);
  return 0;
}

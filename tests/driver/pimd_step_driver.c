/*
 * A driver as a user of the reference library writes one (plain C99, the
 * reference's API only), built against the engine with
 *   cc pimd_step_driver.c $(xmd_config --xmd-flags) $(xmd_config --xmd-libs)
 * It runs the reference driver loop (step (B): md_update_nhc_VV3_1_x,
 * md_reset_force_x, trap, pair, cycle energies, V_B, exchange forces,
 * md_update_nhc_VV3_2_x; md_simulation_x.c:615-792, md_pimd_x.c:1297-1677,
 * 1969-2721) for trapped bosons with a Gaussian pair interaction, with the
 * virial estimator after every step (md_pimd_x.c:4072-4556).
 *
 *   pimd_step_driver N P T xi steps dt in.bin out.bin
 * in.bin:  x then v of every bead (reference bead order), 3 N P doubles each
 * out.bin: x, v, f after the steps (3 N P doubles each), then the virial
 *          estimator after each step (steps doubles)
 */
#include <stdio.h>
#include <stdlib.h>

#include "md_pimd_x.h"
#include "md_simulation_x.h"
#include "md_statistics_x.h"
#include "md_util_x.h"

static int full_force(md_pimd_t_x *pimd) {
  int rc = md_reset_force_x(pimd->sim);
  rc |= md_pimd_calc_trap_force_x(pimd, md_harmonic_trap_force_x);
  rc |= md_pimd_calc_pair_force_x(pimd, md_gaussian_force_x);
  rc |= md_pimd_fast_fill_ENk_x(pimd, 0);
  rc |= md_pimd_fill_VB_x(pimd);
  rc |= md_pimd_fast_fill_force_VB_x(pimd, 0);
  return rc;
}

int main(int argc, char **argv) {
  if (argc != 9) {
    fprintf(stderr, "usage: %s N P T xi steps dt in.bin out.bin\n", argv[0]);
    return 2;
  }
  const int N = atoi(argv[1]), P = atoi(argv[2]), steps = atoi(argv[5]);
  const double T = atof(argv[3]), xi = atof(argv[4]), dt = atof(argv[6]);
  const int nb = N * P;
  md_init_config_x cfg = {1, 1, 0, 1, ""};
  if (md_init_setup_x(NULL, &cfg) != 0) return 1;
  md_set_pimd_mode_x(1);
  md_num_t_x zero[3] = {0, 0, 0}, L[3] = {10, 10, 10};
  md_set_harmonic_trap_frequency_x(1.0);
  md_set_harmonic_trap_center_x(zero);
  md_set_gaussian_force_strength_x(3.0);
  md_set_guassian_force_range_x(0.5);
  md_simulation_t_x *sim = md_alloc_simulation_x(nb, 1, 0, MD_ND_RECT_BOX_X);
  if (!sim) return 1;
  md_init_nd_rect_box_x(sim->box, L);
  md_pimd_t_x *pimd = md_alloc_pimd_x(sim, N, P, xi, T);
  if (!pimd) return 1;
  md_init_particle_mass_x(sim, 1.0, 0, nb);
  double *buf = malloc(sizeof(double) * 6 * nb);
  FILE *in = fopen(argv[7], "rb");
  if (!in || fread(buf, sizeof(double), 6 * (size_t)nb, in) != 6 * (size_t)nb) return 1;
  fclose(in);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < 3; ++d) {
      sim->particles[i].x[d] = buf[3 * i + d];
      sim->particles[i].v[d] = buf[3 * nb + 3 * i + d];
      sim->particles[i].f[d] = 0;
    }
  md_init_nhc_x(sim, md_get_quick_Q_x(T, pimd->omegaP), 0, sim->Nf, 0);
  double *ev = malloc(sizeof(double) * (steps > 0 ? steps : 1));
  md_stats_t_x *st = md_alloc_stats_x(1);
  if (full_force(pimd)) return 1;
  for (int s = 0; s < steps; ++s) {
    if (md_update_nhc_VV3_1_x(sim, dt) || full_force(pimd) || md_update_nhc_VV3_2_x(sim, dt)) return 1;
    if (md_pimd_fast_fill_ENk2_x(pimd, 0)) return 1;
    st->es[0] = 0;
    if (md_pimd_calc_vi_virial_energy_x(pimd, st, 0)) return 1;
    ev[s] = st->es[0];
  }
  if (!md_simulation_sync_host_x(sim, 1)) return 1;
  FILE *out = fopen(argv[8], "wb");
  if (!out) return 1;
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < nb; ++i)
      for (int d = 0; d < 3; ++d) {
        const md_particle_t_x *q = &sim->particles[i];
        double v = k == 0 ? q->x[d] : (k == 1 ? q->v[d] : q->f[d]);
        fwrite(&v, sizeof v, 1, out);
      }
  fwrite(ev, sizeof(double), (size_t)steps, out);
  fclose(out);
  printf("driver: N=%d P=%d %d steps done\n", N, P, steps);
  return 0;
}

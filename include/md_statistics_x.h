/*
 * Block-averaged statistics and energy callbacks (drop-in for the
 * reference's md_statistics_x.h:7-91, nocl build, hot-path subset).
 */
#ifndef MD_STATISTICS_X_H
#define MD_STATISTICS_X_H

#include "md_particle_x.h"
#include "md_simulation_x.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MD_MAX_STATS_COUNT 10000

typedef struct {
  int N;            /* number of blocks (the last one is open) */
  int count;        /* samples in the open block */
  md_num_t_x e;     /* finalised mean of block means */
  md_num_t_x *es;   /* per-block accumulators / means */
  md_num_t_x T;
  md_num_t_x *Ts;
  int points;
  md_num_t_x *fx;
  md_num_t_x **fxs;
} md_stats_t_x;

/* q vectors of a spherical S(q) grid, one representative per |q| shell, sorted by |q|
 * (md_statistics_x.h:28-34). */
typedef struct {
  int N;
  md_num_t_x *qs;   /* |q|, ascending */
  md_num_t_x *qis;  /* representative (qx, qy, qz) per shell */
} md_sk_sphere_t_x;

md_sk_sphere_t_x *md_alloc_sk_sphere_3d_x(md_num_t_x q0, md_num_t_x qincre, md_num_t_x qmax);

md_stats_t_x *md_stats_sync_host_x(md_stats_t_x *stats);

typedef md_num_t_x (*md_pair_energy_t_x)(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
typedef md_num_t_x (*md_trap_energy_t_x)(md_num_t_x *p1, md_num_t_x *L, md_num_t_x m);

md_stats_t_x *md_alloc_stats_x(int points);
md_stats_t_x *md_update_stats_x(md_stats_t_x *stats);
md_stats_t_x *md_finalize_stats_x(md_stats_t_x *stats);

md_num_t_x md_harmonic_trap_energy_x(md_num_t_x *p1, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_gaussian_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_periodic_gaussian_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_coulomb_3d_ewald_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
/* Ewald energy splits (md_statistics_x.h:57-59, md_statistics_x.c:398-495):
 * _R_x every term but the minimum image's real-space one, _R2_x that plus
 * -erf(alpha r)/r of the minimum image, _NI_x the minimum image's
 * erfc(alpha r)/r alone (Ewald = R + NI). */
md_num_t_x md_coulomb_3d_ewald_energy_R_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_coulomb_3d_ewald_energy_R2_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_coulomb_3d_ewald_energy_NI_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_LJ_periodic_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);
md_num_t_x md_periodic_helium_energy_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, md_num_t_x m);

/* md_statistics_x.h:83 (md_statistics_x.c:855-958): Ts[N-1] += sum m v^2 over
 * the sum of the thermostat groups' degree-of-freedom counts. */
int md_calc_temperature_x(md_simulation_t_x *sim, md_stats_t_x *stats);
int md_calc_kinetic_x(md_simulation_t_x *sim, md_stats_t_x *stats);
int md_calc_nhc_kinetic_x(md_simulation_t_x *sim, md_stats_t_x *stats);

#ifdef __cplusplus
}
#endif

#endif

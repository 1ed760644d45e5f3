/*
 * Interactions, integrator and box helpers (drop-in for the reference's
 * md_simulation_x.h:15-82, nocl build).
 *
 * Interaction callbacks keep the reference's plugin contract
 * (md_simulation_x.h:38-39): a callback accumulates into f (+=) an
 * acceleration (force / m), reading its parameters from the globals below.
 * The GPU engine recognises the callbacks listed here by pointer identity and
 * runs its own HIP kernels for them; any other callback pointer passed to a
 * per-step PIMD entry point is rejected with a -1 return (there is no host
 * fallback path).
 */
#ifndef MD_SIMULATION_X_H
#define MD_SIMULATION_X_H

#include "md_particle_x.h"

#ifdef __cplusplus
extern "C" {
#endif

extern md_num_t_x md_trap_frequency_x;
extern md_num_t_x md_trap_center_x[MD_DIMENSION_X];
extern md_num_t_x md_gaussian_strength_x;
extern md_num_t_x md_gaussian_range_x;
extern md_num_t_x md_coulomb_strength_x;
extern md_num_t_x md_coulomb_truc_x;
extern int md_coulomb_n_sum_x;

typedef void (*md_pair_force_t_x)(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
typedef void (*md_trap_force_t_x)(md_num_t_x *p1, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);

/* Minimum-image helpers (md_simulation_x.c:70-114). */
md_num_t_x md_minimum_image_x(md_num_t_x a, md_num_t_x L);
md_num_t_x md_minimum_image_distance_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L);
md_num_t_x md_minimum_image_distance_2_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *L, int *w1, int *w2);
md_num_t_x md_distance_x(md_num_t_x *p1, md_num_t_x *p2);

/* Harmonic trap (md_simulation_x.c:124-137). */
void md_harmonic_trap_force_x(md_num_t_x *p1, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
void md_set_harmonic_trap_frequency_x(md_num_t_x f);
void md_set_harmonic_trap_center_x(md_num_t_x *center);

/* Gaussian pair interaction (g / (pi s^2)) exp(-r^2/s^2) (md_simulation_x.c:163-188). */
void md_gaussian_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
void md_periodic_gaussian_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
void md_set_gaussian_force_strength_x(md_num_t_x g);
void md_set_guassian_force_range_x(md_num_t_x s);   /* reference spelling */

/* Lennard-Jones 4 (r^-12 - r^-6), minimum image (md_simulation_x.c:116-122). */
void md_LJ_periodic_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);

/* Helium pair potential eps [A e^{-alpha x} - (C6 x^-6 + C8 x^-8 + C10 x^-10) F(x)], x = r / r_m,
 * minimum image (md_simulation_x.c:365-399). */
void md_set_helium_parameters_x(md_num_t_x eps, md_num_t_x A, md_num_t_x alpha, md_num_t_x C6, md_num_t_x C8, md_num_t_x C10, md_num_t_x D, md_num_t_x rm);
void md_periodic_helium_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);

/* 3-D Ewald Coulomb (md_simulation_x.c:206-346). */
void md_set_coulomb_force_strength_x(md_num_t_x g);
void md_set_coulomb_truc_x(md_num_t_x k);
void md_set_coulomb_n_sum_x(int n);
md_num_t_x md_calc_3d_ueg_madelung_x(md_num_t_x *L);
void md_coulomb_3d_ewald_force_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
void md_coulomb_3d_ewald_force_R_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);
void md_coulomb_3d_ewald_force_NI_x(md_num_t_x *p1, md_num_t_x *p2, md_num_t_x *f, md_num_t_x *L, md_num_t_x m);

/* Integrator and box (md_simulation_x.c:421-472, 599-911). */
int md_reset_force_x(md_simulation_t_x *sim);
void md_calc_nhc_force_x(md_simulation_t_x *sim, int f0, int f2, int i, md_num_t_x *res);
int md_update_nhc_VV3_1_x(md_simulation_t_x *sim, md_num_t_x h);
int md_update_nhc_VV3_2_x(md_simulation_t_x *sim, md_num_t_x h);
int md_periodic_boundary_x(md_simulation_t_x *sim);
int md_periodic_image_count_x(md_num_t_x x, md_num_t_x L);
int md_periodic_boundary_count_x(md_simulation_t_x *sim, int *count);
int md_periodic_boundary_recover_x(md_simulation_t_x *sim, int *count);
int md_simulation_rescale_velocity_x(md_simulation_t_x *sim, md_num_t_x T);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Path-integral MD object, exchange recursion and estimators (drop-in for
 * the reference's md_pimd_x.h:10-232, nocl build, per-step-path subset).
 *
 * Bead (l, j), 1-based, lives at particles[MD_PIMD_INDEX_X(l, j, P)]
 * (particle-major), exactly as in the reference.  On the device the engine
 * keeps its own slice-major layout and converts at the sync boundary.
 */
#ifndef MD_PIMD_X_H
#define MD_PIMD_X_H

#include "md_particle_x.h"
#include "md_statistics_x.h"
#include "md_simulation_x.h"
#include "md_util_x.h"
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MD_PIMD_INDEX_X(l, j, P) (((l) - 1) * (P) + (j) - 1)

typedef struct {
  md_simulation_t_x *sim;
  int N, P;
  int ENk_count;
  md_num_t_x **ENk;     /* ENk[n-1][k-1] = E_n^(k), triangular rows */
  md_num_t_x **ENk2;    /* virial cross terms, same shape */
  md_num_t_x **ENk3;
  md_num_t_x *VBN;      /* V_B^(0..N) */
  md_num_t_x vi;        /* exchange parameter xi */
  md_num_t_x omegaP;
  md_num_t_x beta;
  int pcount_in, pcount_ex;
  md_index_pair_t_x *pair_in;
  md_index_pair_t_x *pair_ex;
  int *Eindices;
  md_num_t_x *minE;
  int cut_Pj;
  md_num_t_x *Eint;
  md_num_t_x *Evir;
  int mc_par;
  int pbc;
  int *pbW;
  int worm_index;
  md_num_t_x worm_pos[MD_DIMENSION_X];
  int old_worm_index;
  md_num_t_x old_worm_pos[MD_DIMENSION_X];
} md_pimd_t_x;

md_pimd_t_x *md_pimd_sync_host_x(md_pimd_t_x *pimd, int read_only);

md_pimd_t_x *md_alloc_pimd_x(md_simulation_t_x *sim, int N, int P, md_num_t_x vi, md_num_t_x T);
void md_pimd_change_temperature_x(md_pimd_t_x *pimd, md_num_t_x T);
int md_pimd_next_index_x(int l, int j, int N2, int k, int P);
int md_pimd_prev_index_x(int l, int j, int N2, int k, int P);
int md_pimd_init_particle_uniform_pos_x(md_pimd_t_x *pimd, md_num_t_x *center, md_num_t_x *L, md_num_t_x fluc);

/* Exchange: cycle energies, forward V_B recursion, exchange forces. */
int md_pimd_fill_ENk_x(md_pimd_t_x *pimd, int image);
int md_pimd_fast_fill_ENk_x(md_pimd_t_x *pimd, int image);
int md_pimd_fprint_ENk_x(FILE *out, md_pimd_t_x *pimd);
md_num_t_x md_pimd_xexp_x(md_num_t_x k, md_num_t_x E, md_num_t_x EE, md_num_t_x beta, md_num_t_x vi);
md_num_t_x md_pimd_fast_xexp_x(md_num_t_x l, md_num_t_x k, md_num_t_x E, md_num_t_x EE, md_num_t_x beta, md_num_t_x vi);
int md_pimd_fill_VB_x(md_pimd_t_x *pimd);
int md_pimd_fprint_VBN_x(FILE *out, md_pimd_t_x *pimd);
int md_pimd_fill_force_VB_x(md_pimd_t_x *pimd, int image);
int md_pimd_fast_fill_force_VB_x(md_pimd_t_x *pimd, int image);

/* Single-value helpers of the exchange path, evaluated on the device state
 * (md_pimd_x.h:121,125,128,137,139,156,161,171): one cycle energy E_n^(k) by
 * the explicit ring walk, the recursions' log-sum-exp shifts (forward, with a
 * caller-supplied V_B at a second xi, and backward with a caller-supplied V),
 * one bead's spring gradient (cycle form, and by explicit reference bead
 * indices (l-1)P + (j-1)), and the link-virial cycle terms.  NaN (or a zero
 * gradient) when the device call fails. */
md_num_t_x md_pimd_ENk_x(md_pimd_t_x *pimd, int N2, int k, int image);
md_num_t_x md_pimd_xminE_x(md_pimd_t_x *pimd, int N2);
md_num_t_x md_pimd_xminE2_x(md_pimd_t_x *pimd, int N2, md_num_t_x *VBN2, md_num_t_x vi2);
md_num_t_x md_pimd_fast_xminE_x(md_pimd_t_x *pimd, int u, md_num_t_x *V);
void md_pimd_dENk_x(md_pimd_t_x *pimd, int N2, int k, int l, int j, int image, md_num_t_x *dENk);
void md_pimd_fast_dENk_x(md_pimd_t_x *pimd, int index, int index2, int index3, int image, md_num_t_x *dENk);
md_num_t_x md_pimd_ENk2_x(md_pimd_t_x *pimd, int N2, int k, int image);
md_num_t_x md_pimd_fast_ENk2_x(md_pimd_t_x *pimd, int index, int index2, int index3, int image);

/* Per-slice external and pair forces. */
int md_pimd_calc_trap_force_x(md_pimd_t_x *pimd, md_trap_force_t_x tf);
int md_pimd_calc_pair_force_x(md_pimd_t_x *pimd, md_pair_force_t_x pf);

/* Energy estimators. */
int md_pimd_calc_VBN_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats);
int md_pimd_calc_trap_energy_x(md_pimd_t_x *pimd, md_trap_energy_t_x te, md_stats_t_x *stats);
int md_pimd_calc_pair_energy_x(md_pimd_t_x *pimd, md_pair_energy_t_x pe, md_stats_t_x *stats);
/* Pair forces / energy between two species with the same P (md_pimd_x.h:146-147,
 * md_pimd_x.c:3282-3510): every bead of pimd against every bead of pimd2 in
 * the same slice; device kernels for the Gaussian, Lennard-Jones and helium
 * callbacks. */
int md_pimd_calc_pair_force_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_pair_force_t_x pf);
int md_pimd_calc_pair_energy_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_pair_energy_t_x pe, md_stats_t_x *stats);

/* Per-sample observables into stats->fxs[stats->N-1][0..points), nocl
 * semantics; histograms bit-identical to the reference's.
 *  - density distribution around center (md_pimd_x.h:135, md_pimd_x.c:1826-1967);
 *  - ITCF: 2-D histogram of (x_{l,1}, x_{j,pi}) along dimension 0 over [rmin, rmax)
 *    (md_pimd_x.h:141-142, md_pimd_x.c:2723-2872);
 *  - pair correlation of bead 1 against bead pi, one or two systems, and the in-plane
 *    form inside a window box {x0, x1, y0, y1} (md_pimd_x.h:143-145, md_pimd_x.c:2974-3280);
 *  - 2-D |dx|, |dy| pair histograms (md_pimd_x.h:183-184, md_pimd_x.c:4869-4956);
 *  - S(q) on axes, q = q0 + j qincre, and on the spherical shells of md_alloc_sk_sphere_3d_x
 *    (md_pimd_x.h:185-190, md_pimd_x.c:4958-5450). */
int md_pimd_calc_density_distribution_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, int image,
                                        md_num_t_x *center);
int md_pimd_calc_ITCF_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmin, md_num_t_x rmax, int pi);
int md_pimd_calc_ITCF_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x rmin,
                          md_num_t_x rmax, int pi);
int md_pimd_calc_pair_correlation_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                    int pi, int image);
int md_pimd_calc_pair_correlation_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x rmax,
                                      md_num_t_x norm, int pi, int image);
int md_pimd_calc_pair_correlation_2d_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                       md_num_t_x *box);
int md_pimd_calc_2d_pair_correlation_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x rmax, md_num_t_x norm,
                                       int pi, int image);
int md_pimd_calc_2d_pair_correlation_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats,
                                         md_num_t_x rmax, md_num_t_x norm, int pi, int image);
int md_pimd_calc_Sk_structure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, md_num_t_x q0, md_num_t_x qincre, int pi);
int md_pimd_calc_Sk_structure_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x q0,
                                  md_num_t_x qincre, int pi);
int md_pimd_calc_Sk_structure_full_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, md_num_t_x q0,
                                       md_num_t_x qincre, int pi);
int md_pimd_calc_sphere_Sk_structure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int pi, md_sk_sphere_t_x *sk_sphere);
int md_pimd_calc_sphere_Sk_structure_full_2_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_stats_t_x *stats, int pi,
                                              md_sk_sphere_t_x *sk_sphere);
int md_pimd_calc_sphere_Sk_structure_3_x(md_pimd_t_x *pimd, md_pimd_t_x *pimd2, md_pimd_t_x *pimd3,
                                         md_stats_t_x *stats, int pi, md_sk_sphere_t_x *sk_sphere);

int md_pimd_calc_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats);
int md_pimd_calc_centroid_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int fc);
int md_pimd_fill_ENk2_x(md_pimd_t_x *pimd, int image);
int md_pimd_fast_fill_ENk2_x(md_pimd_t_x *pimd, int image);
int md_pimd_fprint_ENk2_x(FILE *out, md_pimd_t_x *pimd);
int md_pimd_calc_vi_virial_energy_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int image);
/* Virial pressure estimator (md_pimd_x.h:191, md_pimd_x.c:5452-5517, pbc = 0); needs a box. */
int md_pimd_calc_vi_virial_pressure_x(md_pimd_t_x *pimd, md_stats_t_x *stats, int image);
int md_pimd_calc_vi_sign_x(md_pimd_t_x *pimd, md_stats_t_x *e_s, md_stats_t_x *s_s, md_num_t_x vi2);

/* Statistics helpers that live in md_pimd_x.c in the reference. */
int md_reset_stats_x(md_stats_t_x *stats, md_num_t_x e);
int md_stats_add_to_x(md_stats_t_x *stats, md_stats_t_x *stats2);
int md_stats_copy_to_x(md_stats_t_x *stats, md_stats_t_x *stats2);
int md_stats_mult_to_x(md_stats_t_x *stats, md_stats_t_x *stats2);
int md_reset_stats_f_x(md_stats_t_x *stats, md_num_t_x e);
int md_stats_add_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2);
int md_stats_copy_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2);
int md_stats_mult_to_f_x(md_stats_t_x *stats, md_stats_t_x *stats2);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Engine extensions of the MI355X PIMD engine.  None of these has a
 * counterpart in the reference; they are additive entry points for drivers
 * that want the fused, graph-captured MD step, run-time knobs, or access to
 * internal exchange quantities for testing.  All follow the reference's
 * conventions: int status 0 = success, -1 = failure.
 */
#ifndef MD_HIP_X_H
#define MD_HIP_X_H

#include "md_pimd_x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Chain length M integrated by md_update_nhc_VV3_1_x / _2_x (even,
 * 2 <= M <= MD_NHC_LENGTH_X; default MD_NHC_LENGTH_X).  Applies to objects
 * uploaded after the call. */
int md_set_nhc_length_x(int M);
int md_get_nhc_length_x(void);

/* Waits for all queued device work of the engine; -1 if the exchange
 * recursion of any object reported a timed-out hand-off since its last
 * check (every object with a recursion launched since then is checked; a
 * reported error is cleared). */
int md_hip_synchronize_x(void);

/* Test hook for the error path above: the next `count` forward recursion
 * solves of this object (N > 128, where the solve hands blocks between
 * workgroups) give up their first hand-off wait at once, as a timed-out
 * wait does.  Every kernel still completes; the affected results are
 * invalid and the next status check reports -1. */
int md_pimd_inject_recursion_timeouts_x(md_pimd_t_x *pimd, int count);

/* Test hook for the exact path of the exchange recursion: each 64-position
 * block is solved in the frame of its single-particle chain, where every
 * scaled value is >= 1, and accepted only while those values (and the
 * look-ahead sums into the next block) stay at or below `limit` (default
 * 1e250, the frame's safe range); a block or row beyond it is solved or
 * folded again exactly.  A limit below 1 sends every block through the exact
 * path; moderate limits mix the two.  Applies to recursions launched or
 * graphs captured afterwards, for every object; limit <= 0 restores the
 * default. */
int md_set_recursion_frame_limit_x(double limit);

/* Cutoff x_cut of both Ewald series in units of the split alpha: real space
 * within x_cut / alpha, k space within pi |G| / alpha <= x_cut (default 5.25:
 * omitted terms below erfc(5.25) = 1.4e-13 and exp(-5.25^2) = 1.1e-12 of the
 * leading ones; 6 brings the sum to the rounding floor for 20 % more work). */
int md_set_ewald_real_cutoff_x(md_num_t_x x_cut);

/* Ewald split used for md_coulomb_3d_ewald_force_x / _energy_x.  The Ewald
 * sum does not depend on the splitting parameter once its real- and
 * reciprocal-space series are converged; when the reference parameters
 * (md_coulomb_truc_x, md_coulomb_n_sum_x) are converged (first omitted k
 * shell below exp(-30), real-space image cube beyond erfc(6)), the engine may
 * evaluate the same sum with its own alpha' (real space within x_cut/alpha',
 * k in the ball pi|G|/alpha' <= x_cut).  alpha_L = 0: always the reference's
 * alpha and k cube; alpha_L > 0: alpha' = alpha_L / min(L) when allowed;
 * alpha_L < 0 (default): alpha' chosen by the engine's cost model.  The
 * _R_x / _NI_x splits always use the reference's alpha. */
int md_set_ewald_split_x(md_num_t_x alpha_L);
/* The split md_coulomb_3d_ewald_force_x would use for this object: alpha,
 * real-space cutoff and the number of half-space k vectors (any may be NULL). */
int md_pimd_get_ewald_plan_x(md_pimd_t_x *pimd, md_num_t_x *alpha, md_num_t_x *r_cut, int *n_k);

/* Lossless checkpoint / restart of a simulation's dynamical state: the x,
 * v, f and nhc blocks in the layout of md_fprint_particle_pos_x / _vel_x /
 * _force_x / md_fprint_nhcs_x (md_util_x.c:838-915), every number a C99
 * hexadecimal float ("%a"), so a run resumes bit for bit and the reference's
 * own md_fread_* readers parse the file.  A last line "rng <seed> <draws>
 * <cache flag> <cache>" records the state of md_random_uniform_x /
 * md_random_gaussian_x (draws made through the library since the last
 * md_set_seed_x); reading restores it.  Reading hands authority back to the
 * host mirror (the next device call uploads it). */
int md_checkpoint_write_x(FILE *out, md_simulation_t_x *sim);
int md_checkpoint_read_x(FILE *in, md_simulation_t_x *sim);

/* Fermion-sign reweighting over a grid of exchange parameters xi' (the xi
 * grid used to extrapolate to fermions): factor[i] = sign(N) exp(-beta
 * V'(N) + beta V_B(N)) with V' the forward recursion at |xi'_i| and the sign
 * tracked through it -- the factor md_pimd_calc_vi_sign_x (md_pimd_x.c:4665-4705)
 * multiplies into its two statistics, for n <= 512 values in one launch
 * (one workgroup per xi'). */
int md_pimd_calc_vi_sign_grid_x(md_pimd_t_x *pimd, int n, const md_num_t_x *vi2, md_num_t_x *factor);

/* Allocation-free variant of md_pimd_fill_VB_x.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Input d_ENk: the cycle-energy triangle in the reference's row
 * layout (E_n^(k) at (n-1)n/2 + k-1, N(N+1)/2 doubles) in device memory;
 * output d_VBN: V_B^(0..N) (N+1 doubles, device memory) and, when d_Vt is not
 * NULL, the backward potentials Vt^(0..N) (device memory).  1 <= N <= 1024,
 * beta > 0, vi >= 0; larger N is rejected with an error.  The working memory
 * (6.2 MiB: the column copy of the cycle energies, the solve state and its
 * coefficients) is a static array in the library's device code that every
 * call on the same GPU shares: calls that overlap on one GPU race.  The call
 * allocates nothing (no hipMalloc, hipFree or allocating library routine);
 * it needs the library set up (md_init_setup_x or any earlier engine call)
 * and is queued on the engine's main stream, ordered after earlier engine
 * calls; the caller makes d_ENk ready before the call and reads the outputs
 * after md_hip_synchronize_x, which also reports a timed-out hand-off of the
 * solve (-1). */
int md_pimd_fill_VB_noalloc_x(int N, md_num_t_x beta, md_num_t_x vi, const md_num_t_x *d_ENk, md_num_t_x *d_VBN,
                              md_num_t_x *d_Vt);

/* Internal exchange quantities after md_pimd_fill_VB_x (tests, diagnostics):
 * backward potentials Vt^(0..N) and the connection-probability matrix
 * G[(l-1)*N + (j-1)] = P(bead P of l links to bead 1 of j). */
int md_pimd_get_backward_potentials_x(md_pimd_t_x *pimd, md_num_t_x *Vt);
int md_pimd_get_connection_matrix_x(md_pimd_t_x *pimd, md_num_t_x *G);

/* One MD step of the reference driver loop (velocity Verlet with NHC, force
 * reset, trap, pair, cycle energies, V_B and exchange forces), fused on the
 * device.  The exchange branch runs on its own stream, concurrently with the
 * pair forces. */
typedef struct {
  md_trap_force_t_x trap; /* md_harmonic_trap_force_x or NULL */
  md_pair_force_t_x pair; /* a supported pair callback or NULL */
  md_num_t_x dt;
  int pbc;       /* wrap positions into the box after the drift */
  int image;     /* minimum-image springs (periodic systems) */
  int use_graph; /* replay a captured hipGraph of the step */
} md_pimd_step_config_x;

/* Forces at the current positions (step stages B3-B6). */
int md_pimd_compute_forces_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg);
/* nsteps full steps (B1-B7).  Eager (use_graph == 0 or nsteps == 1): returns
 * when the work is queued.  Graph (use_graph != 0): the step (four
 * consecutive steps in calls of 32 steps or more) is captured once per call
 * into one graph and replayed (steps left over run eagerly); the executable
 * graph is kept with the object and updated in place by its next call
 * (instantiated again only when the graph's shape changed); returns when the
 * steps are done, -1 if the exchange recursion reported an error. */
int md_pimd_run_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, int nsteps);

/* nsteps steps of n independent systems (replicas of uneven N, P, xi and
 * interaction, each with its own config), one launch per step stage over a
 * packed table of all replicas, captured into one hipGraph per batch step.
 * Every replica's trajectory is bit-identical to running it alone through
 * md_pimd_run_steps_x.  Interaction parameters are the process
 * globals at the time of the call, as for every other entry point.  Returns
 * when the work is done (-1 if any replica's recursion reported an error). */
int md_pimd_run_batch_x(md_pimd_t_x *const *pimds, const md_pimd_step_config_x *cfgs, int n, int nsteps);

/* Learned stand-in for the reciprocal-space Ewald term (no counterpart in
 * the reference).  A per-particle network on symmetry-function descriptors
 * of each particle's neighbourhood in its bead slice (64 radial + 32
 * angular, minimum image, cutoff r_cut <= half the box) predicts that
 * particle's share e_i of its slice's reciprocal energy (the exact split
 * E_rec = sum_i e_i, e_i = sum_k A_k (Re(e^{ik.x_i} S_k*) - 1), of the Ewald
 * split md_coulomb_3d_ewald_force_x uses); forces are -dE/dx through the
 * network and the descriptors.  96 -> 128 -> 128 -> 1 tanh network, BF16
 * matrix-core inference with FP32 accumulation, FP32 training (Adam) on
 * labels from the exact FP64 reciprocal kernels. */
typedef struct md_ewald_surrogate_x md_ewald_surrogate_x;
md_ewald_surrogate_x *md_alloc_ewald_surrogate_x(unsigned seed, md_num_t_x r_cut);
/* Keep the low reciprocal shells |m| <= nlow (G = m / L, ball along the
 * longest side) exact in FP64 and let the network stand in for the rest of
 * the sum only (0, the default: the network stands in for all of it).  The
 * reciprocal force of a dense periodic system is dominated by these few
 * long-wavelength terms, which no cutoff descriptor can see.  Set before
 * training (a change discards the normalisation); the labels, checks, judge
 * and surrogate steps then all use the split. */
int md_ewald_surrogate_set_exact_shells_x(md_ewald_surrogate_x *s, int nlow);
/* iters Adam steps (learning rate lr) on the current configuration's exact
 * labels (the first call also fixes the feature and label normalisation);
 * *loss: the normalised training loss before the last step. */
int md_ewald_surrogate_train_x(md_ewald_surrogate_x *s, md_pimd_t_x *pimd, md_num_t_x lr, int iters,
                               md_num_t_x *loss);
/* Surrogate against the exact FP64 reciprocal term on the current
 * configuration: report[0] mean over slices of (E_pred - E_exact)/N, [1]
 * largest |E_pred - E_exact|/N of a slice (Hartree per particle), [2..4]
 * RMS force error per component, [5..7] RMS exact force per component
 * (Hartree/bohr, slice-energy forces), [8] overall relative RMS force error,
 * [9] mean exact reciprocal energy per particle. */
int md_ewald_surrogate_check_x(md_ewald_surrogate_x *s, md_pimd_t_x *pimd, md_num_t_x *report);
/* Parameters (FP32, 29057 values: W1[96][128], b1[128], W2[128][128], b2,
 * w3[128], b3) and normalisation (mean[96], 1/std[96], label mean, label
 * std): get and/or set (either may be NULL); return the count. */
int md_ewald_surrogate_params_x(md_ewald_surrogate_x *s, float *get, const float *set);
int md_ewald_surrogate_normalization_x(md_ewald_surrogate_x *s, md_num_t_x *get, const md_num_t_x *set);
/* Network prediction (raw descriptors [b][96], energies [b], slice-energy
 * forces [b][3]) and its exact targets ([b], [b][3]: the labels less the
 * exact low shells) on the current configuration, and
 * the training-loss gradient for given per-bead targets, in reference bead
 * order b = (l-1) P + (j-1).  Outputs may be NULL. */
int md_ewald_surrogate_eval_x(md_ewald_surrogate_x *s, md_pimd_t_x *pimd, float *desc, md_num_t_x *e_bead,
                              md_num_t_x *force);
int md_ewald_surrogate_labels_x(md_ewald_surrogate_x *s, md_pimd_t_x *pimd, md_num_t_x *e_bead, md_num_t_x *force);
int md_ewald_surrogate_gradient_x(md_ewald_surrogate_x *s, md_pimd_t_x *pimd, const md_num_t_x *e_bead, float *grad,
                                  md_num_t_x *loss);
/* Online training from the run itself: nsteps exact MD steps (cfg, the
 * UEG/Ewald step of md_pimd_run_steps_x), each followed by `iters` Adam
 * steps on the exact labels of the configuration reached (each bead slice is
 * one configuration).  report (may be NULL): loss of the first and last
 * update. */
int md_pimd_train_surrogate_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, md_ewald_surrogate_x *s,
                                    int nsteps, md_num_t_x lr, int iters, md_num_t_x *report);
/* nsteps MD steps with the surrogate standing in for the reciprocal-space
 * Ewald term (the real-space part, exchange and integrator unchanged;
 * surrogate steps replayed as graphs).  Every check_every-th step is a judge
 * step: the exact FP64 terms drive it and the surrogate is compared with
 * them; when its relative RMS force error exceeds f_threshold or a slice's
 * energy error per particle exceeds e_threshold (Hartree), the next
 * check_every steps run exact, each followed by one online training update
 * (learning rate lr; 0: none), and the surrogate then resumes.  report (8,
 * may be NULL): surrogate steps, judge steps, exact fallback steps, fallback
 * events, mean and largest judge relative force error, mean judge energy
 * error per particle (signed), largest judge |energy error| per particle. */
int md_pimd_run_steps_surrogate_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, md_ewald_surrogate_x *s,
                                  int nsteps, int check_every, md_num_t_x f_threshold, md_num_t_x e_threshold,
                                  md_num_t_x lr, md_num_t_x *report);

/* Autotuner of the Ewald real-space kernel's launch choices, none of which
 * changes a result bit: the image enumeration (face / near / corner; every
 * variant as general as the geometry needs takes the same images in the
 * same order) and the waves per workgroup (1, 2, 4, 8; each wave owns whole
 * tile pairs).  md_pimd_autotune_x times every valid choice on the object's
 * current positions (reps launches each; ms[4 * enumeration + log2 waves],
 * -1 where invalid; choice: the fastest) and keeps it for the object's
 * bucket (N rounded up to a power of two, P, the most specific valid
 * enumeration).  md_set_autotune_x(1): use the kept choices, tuning a bucket
 * on its first use by md_pimd_run_steps_x / md_pimd_run_batch_x (outside any
 * capture); 0 (default): the rule (most specific enumeration, 4 waves).
 * md_autotune_save_x / _load_x write / read the kept choices as text (load
 * returns the count).  md_set_ewald_real_variant_x fixes a choice for
 * diagnostics (-1, -1: automatic). */
int md_set_autotune_x(int on);
int md_pimd_autotune_x(md_pimd_t_x *pimd, int reps, md_num_t_x *ms, int *choice);
int md_autotune_save_x(const char *path);
int md_autotune_load_x(const char *path);
int md_set_ewald_real_variant_x(int img, int wlog);

/* Per-stage device time of nsteps eager steps, averaged per step (ms).  With
 * serial != 0 every stage runs alone on one stream (true kernel times);
 * otherwise the branches overlap as in md_pimd_run_steps_x (stage times are
 * then bracketed only where a stage still runs on the main stream):
 *   [0] NHC first half + wrap   [1] cycle energies   [2] V_B recursions
 *   [3] reset/trap/ring springs [4] pair real space  [5] pair reciprocal
 *   [6] exchange end-bead forces [7] NHC second half [8] whole step (wall)
 *   [9] the persistent recursion solve alone (part of [2]; serial only)
 *   [10] the tabulated real-space Ewald kernel alone, without its partial-
 *        force reduction (part of [4]; serial only; 0 for other pair forces)
 * Each stage is bracketed by events on the stream it runs on. */
#define MD_PIMD_NSTAGES_X 11
int md_pimd_profile_steps_x(md_pimd_t_x *pimd, const md_pimd_step_config_x *cfg, int nsteps, int serial,
                            double *stage_ms);

#ifdef __cplusplus
}
#endif

#endif

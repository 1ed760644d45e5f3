/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64, single-threaded restatement of the reference's per-MD-step
 * PIMD path (xiongyunuo/MD-Lab, nocl build).  It is the checker that the HIP
 * engine is compared against; nothing in the product links or calls it.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it.
 *
 * Every function works on caller-owned arrays in the reference's own layout:
 * beads are orc_particle_t records (same 80-byte layout as md_particle_t_x),
 * bead (l, j) (1-based) at index (l-1)*P + (j-1); thermostat chains are
 * orc_nhc_t records (same layout as md_nhc_t_x with an 8-slot chain).
 * Interaction parameters are passed explicitly instead of through globals.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   - golden vectors written by the compiled reference (tests/golden/), for
 *     pair/trap forces and energies, Ewald, ENk (fast and slow), V_B, exchange
 *     forces (fast and slow), ENk2 (fast nocl orientation and slow), the
 *     estimators, NHC VV3 trajectories, RNG/Maxwell initial states;
 *   - closed forms: V_B at xi=0 and for coincident beads, the UEG Madelung
 *     constant, G doubly stochastic, backward endpoint = forward total,
 *     -dV_B/dx by finite differences, ideal trapped bosons.
 *   orc_vi_sign (fermion sign reweighting) is pinned by the golden vectors
 *   only for xi' in the fixture set.
 */
#ifndef ORC_XMD_H
#define ORC_XMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DIM 3
#define ORC_NHC_SLOTS 8

typedef struct {
  double m;
  double x[ORC_DIM];
  double v[ORC_DIM];
  double f[ORC_DIM];
} orc_particle_t;

typedef struct {
  int f;
  double theta[ORC_NHC_SLOTS];
  double vtheta[ORC_NHC_SLOTS];
  double Q[ORC_NHC_SLOTS];
} orc_nhc_t;

/* Interaction parameters (the reference keeps these in globals,
 * md_simulation_x.c:124-221). */
typedef struct {
  double trap_omega;
  double trap_center[ORC_DIM];
  double gauss_g, gauss_s;
  double coul_q, ewald_alpha;
  int ewald_n;
  double he[8]; /* helium: eps, A, alpha, C6, C8, C10, D, r_m (md_simulation_x.c:365-383) */
} orc_params_t;

enum {
  ORC_PAIR_GAUSSIAN = 0,
  ORC_PAIR_PERIODIC_GAUSSIAN = 1,
  ORC_PAIR_EWALD = 2,
  ORC_PAIR_EWALD_R = 3,
  ORC_PAIR_EWALD_NI = 4,
  ORC_PAIR_LJ = 5,
  ORC_PAIR_HELIUM = 6
};

/* Scalar helpers. */
double orc_minimum_image(double a, double L);
int orc_periodic_image_count(double x, double L);

/* Single-pair callbacks (force accumulates into f, an acceleration). */
void orc_pair_force_one(int kind, const orc_params_t *p, const double *p1, const double *p2,
                        double *f, const double *L, double m);
double orc_pair_energy_one(int kind, const orc_params_t *p, const double *p1, const double *p2,
                           const double *L);
double orc_madelung(const orc_params_t *p, const double *L);

/* Per-slice forces and energies of a PIMD system (N particles, P beads). */
void orc_trap_force(const orc_params_t *p, int N, int P, orc_particle_t *b);
void orc_pair_force(int kind, const orc_params_t *p, int N, int P, const double *L, orc_particle_t *b);
double orc_trap_energy(const orc_params_t *p, int N, int P, const orc_particle_t *b);
double orc_pair_energy(int kind, const orc_params_t *p, int N, int P, const double *L,
                       const orc_particle_t *b);
void orc_pair_force_2(int kind, const orc_params_t *p, int NA, int NB, int P, const double *L, orc_particle_t *a,
                      orc_particle_t *c);
double orc_pair_energy_2(int kind, const orc_params_t *p, int NA, int NB, int P, const double *L,
                         const orc_particle_t *a, const orc_particle_t *c);
void orc_density_distribution(int N, int P, const orc_particle_t *b, const double *L, int image,
                              const double *center, double rmax, int points, double *fx);
void orc_pair_correlation(int N, int P, const orc_particle_t *b, const double *L, int image, double rmax,
                          double norm, int pi, int points, double *fx);
void orc_itcf(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, double rmin, double rmax,
              int pi, int points, int two, double *fx);
void orc_pair_correlation_2(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, const double *L,
                            int image, double rmax, double norm, int pi, int points, double *fx);
void orc_pair_correlation_2d(int N, int P, const orc_particle_t *b, double rmax, double norm, const double *box,
                             int points, double *fx);
void orc_2d_pair_correlation(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b,
                             const double *L, int image, double rmax, double norm, int pi, int points, int same,
                             double *fx);
void orc_sk_structure_2(int NA, int NB, int P, const orc_particle_t *a, const orc_particle_t *b, double q0,
                        double qincre, int pi, int points, int full, double *fx);
void orc_sphere_sk(int variant, int NA, int NB, int NC, int P, const orc_particle_t *a, const orc_particle_t *b,
                   const orc_particle_t *c, int pi, const double *qis, int nq, double *fx);
void orc_sk_structure(int N, int P, const orc_particle_t *b, double q0, double qincre, int pi, int points,
                      double *fx);

/* Exchange.  ENk is a packed triangle: E_n^(k) at (n-1)n/2 + k-1. */
void orc_fast_fill_ENk(int N, int P, double omegaP, int image, const double *L,
                       const orc_particle_t *b, double *ENk, double *Eint);
void orc_fill_ENk_slow(int N, int P, double omegaP, int image, const double *L,
                       const orc_particle_t *b, double *ENk);
void orc_fill_VB(int N, double beta, double vi, const double *ENk, double *VBN);
/* Backward potentials Vt[0..N] and connection matrix G (N*N, row-major),
 * either may be NULL.  Adds -grad V_B / m into b[].f. */
void orc_fast_force_VB(int N, int P, double beta, double vi, double omegaP, int image,
                       const double *L, const double *ENk, const double *VBN,
                       orc_particle_t *b, double *Vt, double *G);
void orc_slow_force_VB(int N, int P, double beta, double vi, double omegaP, int image,
                       const double *L, const double *ENk, const double *VBN, orc_particle_t *b);

/* Estimators (return the value the reference adds into stats->es). */
double orc_VBN_energy(int N, int P, double beta, double vi, const double *ENk, const double *VBN);
/* The exchange path (cycle energies, V_B, backward recursion, G, exchange and
 * ring-spring forces) in 80-bit long double: the extended-precision truth. */
void orc_exchange_forces_ld(int N, int P, double beta, double vi, double omegaP, int image, const double *L,
                            const orc_particle_t *b, double *f_out, double *VBN_out);
void orc_fast_fill_ENk2(int N, int P, double omegaP, int image, const double *L,
                        const orc_particle_t *b, double *ENk2);
void orc_fill_ENk2_slow(int N, int P, double omegaP, int image, const double *L,
                        const orc_particle_t *b, double *ENk2);
double orc_vi_virial_energy(int N, int P, double beta, double vi, int image, const double *L,
                            const double *ENk, const double *ENk2, const double *VBN,
                            const orc_particle_t *b);
double orc_vi_virial_pressure(int N, int P, double beta, double vi, int image, const double *L,
                              const double *ENk, const double *ENk2, const double *VBN,
                              const orc_particle_t *b);
double orc_virial_energy(int N, int P, const double *L, const orc_particle_t *b);
double orc_centroid_virial_energy(int N, int P, double beta, int fc, const orc_particle_t *b);
/* Sign reweighting: returns the factor s the reference multiplies into the
 * estimator accumulators (md_pimd_x.c:4699). */
double orc_vi_sign(int N, double beta, double vi2, const double *ENk, const double *VBN);

/* Integrator.  nb = beads, M = chain length actually integrated (even). */
void orc_nhc_vv3_1(int nb, int Nf, orc_nhc_t *nhc, orc_particle_t *b, double T, double h, int M);
void orc_nhc_vv3_2(int nb, int Nf, orc_nhc_t *nhc, orc_particle_t *b, double T, double h, int M);
void orc_periodic_boundary(int nb, const double *L, orc_particle_t *b);
double orc_kinetic(int nb, const orc_particle_t *b);
double orc_nhc_kinetic(int Nf, const orc_nhc_t *nhc, double T, int M);

/* Seeded initial state (glibc rand(), consumed in the reference's order). */
void orc_set_seed(unsigned int seed);
double orc_random_uniform(double a, double b);
double orc_random_gaussian(void);
void orc_init_maxwell_vel(int nb, orc_particle_t *b, double T);
void orc_init_uniform_pos(int N, int P, const double *center, const double *L, double fluc,
                          orc_particle_t *b);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Core data types of the MI355X PIMD engine.
 *
 * Layout contract: every struct here has the exact field order and types of
 * the reference's single-precision-free ("nocl") build, so that a C driver
 * written against xiongyunuo/MD-Lab compiles and links unchanged.
 * Reference: md_particle_x.h:4-107 (the #ifdef MD_USE_OPENCL_X fields of the
 * reference are not part of the nocl ABI and are absent here; device state
 * lives in side tables owned by the library).
 */
#ifndef MD_PARTICLE_X_H
#define MD_PARTICLE_X_H

#define MD_DIMENSION_X 3
/* Storage length of a Nose-Hoover chain record (md_particle_x.h:5).  The
 * chain length actually integrated is a run-time setting
 * (md_set_nhc_length_x in md_hip_x.h), any even value <= this. */
#define MD_NHC_LENGTH_X 8
#define MD_kB_X 1.0
#define MD_hBar_X 1.0

#include <math.h>

#ifndef M_PI
#define M_PI 3.141592653589793
#endif

#ifdef __GNUC__
#define MD_UNUSED_X(x) UNUSED_##x __attribute__((__unused__))
#else
#define MD_UNUSED_X(x) UNUSED_##x
#endif

typedef double md_num_t_x;

/* One bead: mass, position, velocity and acceleration (force / m). */
typedef struct {
  md_num_t_x m;
  md_num_t_x x[MD_DIMENSION_X];
  md_num_t_x v[MD_DIMENSION_X];
  md_num_t_x f[MD_DIMENSION_X];
} md_particle_t_x;

/* One thermostat chain: f = degrees of freedom it controls. */
typedef struct {
  int f;
  md_num_t_x theta[MD_NHC_LENGTH_X];
  md_num_t_x vtheta[MD_NHC_LENGTH_X];
  md_num_t_x Q[MD_NHC_LENGTH_X];
} md_nhc_t_x;

enum { MD_ND_RECT_BOX_X };

typedef struct {
  int type;
  void *box;
} simulation_box_t_x;

typedef struct {
  md_num_t_x L[MD_DIMENSION_X];
} nd_rect_t_x;

typedef struct {
  int i, j;
} md_index_pair_t_x;

typedef struct {
  int N;                        /* beads (N*P for a PIMD system) */
  md_particle_t_x *particles;   /* host mirror, valid after sync_host */
  int fc;                       /* thermostatted degrees of freedom */
  int Nf;                       /* number of thermostat chains */
  int *fs;
  int *f0s;
  md_nhc_t_x *nhcs;
  simulation_box_t_x *box;
  md_num_t_x T;
  int pcount_in, pcount_ex;
  md_index_pair_t_x *pair_in;
  md_index_pair_t_x *pair_ex;
} md_simulation_t_x;

#endif

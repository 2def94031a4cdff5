/*
 * oracle.h -- CPU restatement of FLITE3D's flow-solver hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only to check
 * results against (or, for cpu_baseline, as the timed CPU comparator).  The product path
 * (flite3d_amd, libflite3d_amd.so) never links, loads or calls it.
 *
 * It restates, routine by routine and in the reference's loop order, the
 * GridSolver module of SolverSource/mgnsg3d_def_par.F90 and the cycle/monitor
 * logic of SolverSource/mgnsg3d_mg_par.F90; each function cites the lines it
 * follows.  It consumes the same flite_params / flite_level_view the
 * C-ABI takes (include/flite3d_amd.h).
 *
 * Parity status (how each part is pinned) is recorded in oracle.c's header.
 */
#ifndef FLITE_ORACLE_H
#define FLITE_ORACLE_H

#include "../include/flite3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const flite_params *p, int n_levels);
int or_level_upload(or_ctx *c, int level, const flite_level_view *v);
int or_level_geometry(or_ctx *c, int level, const flite_level_view *v);
void or_destroy(or_ctx *c);
const char *or_last_error(void);

int or_field_get(or_ctx *c, int level, int field, double *out, int ld);
int or_field_set(or_ctx *c, int level, int field, const double *in, int ld);

int or_apply(or_ctx *c, int level, int op);
/* OpenMP threads of the side-flux loops (the timed CPU comparator); the
 * results do not depend on the count.  1 by default. */
int or_set_threads(or_ctx *c, int threads);
/* sum over the terms an operation adds (RHS, VISCOUS, SA_DIFFUSION) of their
 * magnitudes, per node and component: out(N, 6) */
int or_magnitudes(or_ctx *c, int level, int op, double *out, int ld);
int or_init_freestream(or_ctx *c);
int or_init_from_state(or_ctx *c);
int or_do_time_iterations(or_ctx *c, int level, int steps, int calc_diss,
                          int calc_turb, int iteration);
int or_restrict(or_ctx *c, int fine_level, int calc_turb);
int or_prolongate(or_ctx *c, int coarse_level);
int or_map_u_fine_to_coarse(or_ctx *c, int fine_level);
int or_cycle(or_ctx *c, int cycle_number);
/* end of a physical time step: uold2 <- uold <- u on level 1 (MG:889-890) */
int or_advance_physical_time(or_ctx *c);
int or_monitor(or_ctx *c, double res[7], double maxres[7], int maxidx[7],
               double forces[9]);

/* one fine-level HLLC flux evaluation between two given states (for the
 * closed-form flux tests): writes f[5] for side weights w[3]; order 1. */
int or_hllc_flux(const flite_params *p, const double ul[5], double pl,
                 const double ur[5], double pr, const double w[3], double f[5]);

/* one interior side of makeRHS's HLLC loop (DEF:1794-2100) at either order,
 * with the ends' nodeHelpArray rows given: g1, g2 = 15 gradient components
 * (x of rho,u,v,w,E, then y, then z) and the LocalMinMax bounds (5 minima,
 * 5 maxima, relative to the node); dx = x2 - x1; f[5] */
int or_hllc_side(const flite_params *p, int order, int autoclip, int nsig,
                 const double ul[5], double pl, const double ur[5], double pr,
                 const double w[3], const double dx[3], const double g1[25],
                 const double g2[25], double f[5]);
/* the LED limiter of the reconstruction (DEF:1852-1889) */
double or_led(double a, double bmin, double bmax, int nsig);

/* interior flux of fluxType 1-3 (DEF:1664-1792, LLF and Roe patched as
 * documented in oracle.c) between two given states, side weights w[3] */
int or_simple_flux(const flite_params *p, int flux_type, const double ul[5], double pl,
                   const double ur[5], double pr, const double w[3], double f[5]);

/* farfield Roe flux of one boundary-side end (DEF:2258-2446), unit normal n */
int or_farfield_flux(const flite_params *p, int indicator, const double s[5], double pr,
                     const double n[3], double f[5]);

#ifdef __cplusplus
}
#endif
#endif

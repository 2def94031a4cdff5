/*
 * oracle.c -- CPU restatement (FP64, scalar) of FLITE3D's flow-solver hot
 * path.  TEST INFRASTRUCTURE: see oracle.h for who may use it.
 *
 * Reference files (cited as DEF:line and MG:line):
 *   DEF = SolverSource/mgnsg3d_def_par.F90
 *   MG  = SolverSource/mgnsg3d_mg_par.F90
 *
 * The reference is Fortran built against MPI with a patch set it needs to
 * compile and run here, so it is not built in this repository; this file is
 * pinned instead by the checks in tests/test_oracle_*.py, each a fact the
 * reference's algorithm fixes independently of how this file writes it:
 *   - freestream preservation of the full residual (closure of the side
 *     coefficients, DEF:14430-14568) for HLLC-1/2 and the viscous terms;
 *   - conservation: sum over nodes of the interior scatter is zero;
 *   - HLLC consistency F(U,U,n) = F(U).n, supersonic upwinding, mirror
 *     symmetry; the farfield Roe flux against the textbook Roe solver;
 *   - MUSCL: a linear field reconstructs both faces to the side midpoint
 *     (side flux = physical flux there); minmod and the clip to exactly
 *     half the clip factor times the side length; LED: slope 1 at zero,
 *     the limit bmax/2, c <= min(a, bmax/2), c(bmax/2) = 2^(-1/nsig) bmax/2;
 *   - Green-Gauss exactness for linear fields at every node;
 *   - LocalMinMax, time steps, pressure and the SA time steps by brute force
 *     on small meshes (independent numpy restatements);
 *   - the update: density relaxation by exactly maxIncrementFactor, and one
 *     smoothing step applying the Runge-Kutta stage polynomial
 *     1 - a3 z + a2 a3 z^2 - a1 a2 a3 z^3 to a small perturbation of a
 *     uniform flow (3 and 5 stages);
 *   - forces: lift and drag of a uniform overpressure and of a uniform
 *     shear on a flat wall patch, from the patch's area alone;
 *   - viscosity: Sutherland's law, the SA eddy viscosity (fv1 = 1/2 at
 *     chi = cv1, mu_t -> rho nu~/Re), zero momentum and the closed-form
 *     viscous work of a uniform stress at every node, the turbulent heat
 *     conduction's Prandtl numbers by the ratio of SA to laminar terms;
 *   - SA source: the log-layer balance the model's cw1 is defined by, the
 *     destruction fw(r) of the published model at r != 1 (cw2, cw3, fv2,
 *     the clip at 100), the production and trip term away from walls, zero
 *     for nu~ = 0; SA diffusion exact for linear nu~; SA convection upwind
 *     (no upstream influence);
 *   - MG: restriction conserves sum V u; prolongation injects u_H - R u_h
 *     scaled exactly by its two relaxation factors; a converged fine
 *     solution gives a zero coarse correction; a converged physical step of
 *     multigrid cycles solves the fine grid's BDF equation;
 *   - central / LLF / Roe interior fluxes: consistency F(U,U) = F(U).n,
 *     the Roe flux against the textbook Roe solver, upwinding of a contact;
 *   - JST dissipation: zero for uniform states, conservative (sums to zero
 *     over the nodes), the scalar and matrix forms on hand-computed sides,
 *     the fourth-difference coefficient dim(d4, d2).
 * Branches not reached by those checks (engine faces, inviscid-wall
 * indicator 9, dissipation schemes 1 and 3, k-omega/SST) are not restated:
 * the calls that would need them return FLITE_E_UNSUPPORTED.
 *
 * Known reference defects and how they are mirrored:
 *   - coarse-level coordinates are never filled (DEF:14002,14011,14071):
 *     taken as zero, so HLLC-2 reconstruction vanishes on coarse levels;
 *   - nullifyOuterBoundary reads an uninitialised local ivd (DEF:4259):
 *     treated as a no-op;
 *   - brp%numberOfIONodes is never set (DEF:5132): treated as zero;
 *   - the laminarViscosity = 0.1*Re guard in the SA source (DEF:9328-9330)
 *     is mirrored literally.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXLEV 8

static __thread char g_err[256];

static int fail(int code, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

const char *or_last_error(void) { return g_err; }

typedef struct {
  int gnum; /* gridNumber, 1-based */
  int n, ne, nbs, nb, nfine;
  int *e1, *e2;
  double *w; /* w[c*ne+e] */
  int *b1, *b2, *bi;
  double *wb; /* wb[c*nbs+s] */
  double *vol, *xyz, *wd;
  int reg[21];
  int *nia;
  double *nrm, *warea;
  int *f2c;
  double *rw;
  /* state, component-major [c*n+i] */
  double *u, *uprev, *rhs, *diss, *src;
  double *uold, *uold2; /* solutions at the last two physical times (unsteady) */
  double *p, *dt, *dtt, *mu, *vort, *divg, *tdiff;
  double *ws; /* wallStress(Nb,4): ws[c*nb+i] */
  /* moving mesh (NULL on a fixed mesh): GCLWeightsArray(E), brp%GCLWeightsArray
   * (Eb,2) as gcl_b[end*nbs+s], nodeVolumeP/P2(N) and coordinateMovement(N,3)
   * on level 1 */
  double *gcl, *gclb, *volp, *volp2, *cmv;
  int has_src;
  int uploaded;
} orl;

struct or_ctx {
  flite_params P;
  int nl;
  orl L[MAXLEV];
  double *help; /* nodeHelpArray(Nfine,25), shared by all levels (DEF:14214-14231) */
  int nh;
  int pending; /* a level's geometry was replaced: the step's preamble is still due */
  /* or_magnitudes: when set, the side loops of the operation being run add
   * |contribution| per node and component here ([c*n+i], 6 components) */
  double *mag;
  int threads; /* OpenMP threads of the side-flux and node loops (or_set_threads; 1 by default) */
};

static void settle(or_ctx *C);

/* |v| into the magnitude accumulator of or_magnitudes, when active */
static inline void MAG(or_ctx *C, const orl *L, int c, int i, double v) {
  if (C->mag) C->mag[(size_t)c * L->n + i] += fabs(v);
}

#define U(L, c, i) ((L)->u[(size_t)(c) * (L)->n + (i)])
#define UP(L, c, i) ((L)->uprev[(size_t)(c) * (L)->n + (i)])
#define RHS(L, c, i) ((L)->rhs[(size_t)(c) * (L)->n + (i)])
#define H(C, c, i) ((C)->help[(size_t)(c) * (C)->nh + (i)])
#define A6(a, L, c, i) ((a)[(size_t)(c) * (L)->n + (i)])

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* nodeIndicatorRegister(k) for k in -20..0 */
static int REG(const orl *L, int k) { return L->reg[k + 20]; }
/* iterate boundary nodes of one indicator range (1-based ib in ist..ien) */
#define FOR_RANGE(L, kstart, ib, ip)                                          \
  for (int ib = REG(L, kstart); ib <= REG(L, (kstart) + 1); ++ib)             \
    for (int ip = (L)->nia[ib - 1], _once = 1; _once; _once = 0)

/* ------------------------------------------------------------------------ */

or_ctx *or_create(const flite_params *p, int n_levels) {
  if (!p || n_levels < 1 || n_levels > MAXLEV) {
    fail(FLITE_E_ARG, "bad arguments");
    return NULL;
  }
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->P = *p;
  c->nl = n_levels;
  c->threads = 1;
  return c;
}

static void free_level(orl *L) {
  void *ptrs[] = {L->e1, L->e2, L->w, L->b1, L->b2, L->bi, L->wb, L->vol,
                  L->xyz, L->wd, L->nia, L->nrm, L->warea, L->f2c, L->rw,
                  L->u, L->uprev, L->rhs, L->diss, L->src, L->p, L->dt,
                  L->dtt, L->mu, L->vort, L->divg, L->tdiff, L->ws, L->uold, L->uold2,
                  L->gcl, L->gclb, L->volp, L->volp2, L->cmv};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; ++k) free(ptrs[k]);
  memset(L, 0, sizeof *L);
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  for (int l = 0; l < MAXLEV; ++l) free_level(&c->L[l]);
  free(c->help);
  free(c);
}

/* makeBoundaryNormals (DEF:15394-15460) and makeSurfaceArea (DEF:15462-15540) */
static void make_normals_area(orl *L) {
  for (int s = 0; s < L->nbs; ++s) {
    int i1 = L->b1[s], i2 = L->b2[s];
    for (int c = 0; c < 3; ++c) {
      L->nrm[(size_t)c * L->n + i1] -= L->wb[(size_t)c * L->nbs + s];
      L->nrm[(size_t)c * L->n + i2] -= L->wb[(size_t)c * L->nbs + s];
    }
  }
  for (int i = 0; i < L->nb; ++i) {
    double x = L->nrm[i], y = L->nrm[L->n + i], z = L->nrm[2 * (size_t)L->n + i];
    double nn = sqrt(x * x + y * y + z * z);
    for (int c = 0; c < 3; ++c)
      L->nrm[(size_t)c * L->n + i] = nn > 1.0e-10 ? L->nrm[(size_t)c * L->n + i] / nn : 0.0;
  }
  if (L->gnum == 1) {
    for (int s = 0; s < L->nbs; ++s) {
      double wx = L->wb[s], wy = L->wb[L->nbs + s], wz = L->wb[2 * (size_t)L->nbs + s];
      double a = sqrt(wx * wx + wy * wy + wz * wz);
      L->warea[L->b1[s]] += 4.0 * a;
      L->warea[L->b2[s]] += 4.0 * a;
    }
  }
}

int or_level_upload(or_ctx *c, int level, const flite_level_view *v) {
  if (!c || !v || level < 1 || level > c->nl) return fail(FLITE_E_ARG, "bad level");
  orl *L = &c->L[level - 1];
  free_level(L);
  int b = v->index_base;
  L->gnum = level;
  L->n = v->n_nodes;
  L->ne = v->n_sides;
  L->nbs = v->n_bsides;
  L->nb = v->n_bnodes;
  int n = L->n, ne = L->ne, nbs = L->nbs;
  L->e1 = (int *)malloc(sizeof(int) * (ne ? ne : 1));
  L->e2 = (int *)malloc(sizeof(int) * (ne ? ne : 1));
  L->w = zalloc(3 * (size_t)ne);
  for (int e = 0; e < ne; ++e) {
    L->e1[e] = v->side_index[e] - b;
    L->e2[e] = v->side_index[(size_t)ne + e] - b;
    if (L->e1[e] < 0 || L->e1[e] >= n || L->e2[e] < 0 || L->e2[e] >= n)
      return fail(FLITE_E_ARG, "side index out of range");
  }
  memcpy(L->w, v->side_weights, sizeof(double) * 3 * (size_t)ne);
  L->b1 = (int *)malloc(sizeof(int) * (nbs ? nbs : 1));
  L->b2 = (int *)malloc(sizeof(int) * (nbs ? nbs : 1));
  L->bi = (int *)malloc(sizeof(int) * (nbs ? nbs : 1));
  L->wb = zalloc(3 * (size_t)nbs);
  for (int s = 0; s < nbs; ++s) {
    L->b1[s] = v->bside_index[s] - b;
    L->b2[s] = v->bside_index[(size_t)nbs + s] - b;
    L->bi[s] = v->bside_index[2 * (size_t)nbs + s];
    if (L->b1[s] < 0 || L->b1[s] >= n || L->b2[s] < 0 || L->b2[s] >= n)
      return fail(FLITE_E_ARG, "boundary side index out of range");
  }
  memcpy(L->wb, v->bside_weights, sizeof(double) * 3 * (size_t)nbs);
  L->vol = zalloc(n);
  memcpy(L->vol, v->node_volume, sizeof(double) * n);
  L->xyz = zalloc(3 * (size_t)n);
  if (v->coordinates && level == 1) memcpy(L->xyz, v->coordinates, sizeof(double) * 3 * (size_t)n);
  L->wd = zalloc(n);
  memcpy(L->wd, v->wall_distance, sizeof(double) * n);
  memcpy(L->reg, v->node_indicator_register, sizeof(int) * 21);
  L->nia = (int *)malloc(sizeof(int) * (L->nb ? L->nb : 1));
  for (int i = 0; i < L->nb; ++i) L->nia[i] = v->node_indicator_array[i] - b;
  L->nrm = zalloc(3 * (size_t)n);
  L->warea = zalloc(n);
  make_normals_area(L);
  if (level > 1) {
    L->nfine = v->n_fine;
    L->f2c = (int *)malloc(sizeof(int) * L->nfine);
    L->rw = zalloc(L->nfine);
    for (int i = 0; i < L->nfine; ++i) {
      L->f2c[i] = v->prolongation[i] - b;
      if (L->f2c[i] < 0 || L->f2c[i] >= n) return fail(FLITE_E_ARG, "prolongation index out of range");
    }
    memcpy(L->rw, v->restriction, sizeof(double) * L->nfine);
  }
  L->u = zalloc(6 * (size_t)n);
  L->uprev = zalloc(6 * (size_t)n);
  L->rhs = zalloc(6 * (size_t)n);
  L->diss = zalloc(5 * (size_t)n);
  L->uold = zalloc(6 * (size_t)n);
  L->uold2 = zalloc(6 * (size_t)n);
  L->has_src = level > 1; /* sourceTerm is nullified on grid 1 (DEF:14212-14213) */
  L->src = zalloc(6 * (size_t)n);
  L->p = zalloc(n);
  L->dt = zalloc(n);
  L->dtt = zalloc(n);
  L->mu = zalloc(n);
  L->vort = zalloc(n);
  L->divg = zalloc(n);
  L->tdiff = zalloc(n);
  L->ws = zalloc(4 * (size_t)(L->nb ? L->nb : 1));
  if (level == 1) {
    free(c->help);
    c->nh = n;
    c->help = zalloc(25 * (size_t)n);
  }
  if (c->P.flow_type != 0) { /* moving-mesh arrays: unsteady runs only */
    if (v->side_gcl) { L->gcl = zalloc(ne); memcpy(L->gcl, v->side_gcl, sizeof(double) * ne); }
    if (v->bside_gcl) { L->gclb = zalloc(2 * (size_t)nbs); memcpy(L->gclb, v->bside_gcl, sizeof(double) * 2 * nbs); }
    if (level == 1 && v->node_volume_p) { L->volp = zalloc(n); memcpy(L->volp, v->node_volume_p, sizeof(double) * n); }
    if (level == 1 && v->node_volume_p2) { L->volp2 = zalloc(n); memcpy(L->volp2, v->node_volume_p2, sizeof(double) * n); }
    if (level == 1 && v->coordinate_movement) {
      L->cmv = zalloc(3 * (size_t)n);
      memcpy(L->cmv, v->coordinate_movement, sizeof(double) * 3 * n);
    }
  }
  L->uploaded = 1;
  return 0;
}

/* readComputationData(COMPINFILE, 1) of a moving-mesh step (MG:525-548):
 * new geometry, the solution arrays kept */
int or_level_geometry(or_ctx *c, int level, const flite_level_view *v) {
  if (!c || !v || level < 1 || level > c->nl) return fail(FLITE_E_ARG, "bad level");
  orl *L = &c->L[level - 1];
  if (!L->uploaded || v->n_nodes != L->n || v->n_sides != L->ne || v->n_bsides != L->nbs || v->n_bnodes != L->nb)
    return fail(FLITE_E_ARG, "the moved mesh must keep the level's counts");
  size_t n6 = 6 * (size_t)L->n;
  double *keep[4] = {zalloc(n6), zalloc(n6), zalloc(n6), zalloc(L->n)};
  memcpy(keep[0], L->u, sizeof(double) * n6);
  memcpy(keep[1], L->uold, sizeof(double) * n6);
  memcpy(keep[2], L->uold2, sizeof(double) * n6);
  memcpy(keep[3], L->mu, sizeof(double) * L->n);
  int rc = or_level_upload(c, level, v);
  if (!rc) {
    memcpy(L->u, keep[0], sizeof(double) * n6);
    memcpy(L->uold, keep[1], sizeof(double) * n6);
    memcpy(L->uold2, keep[2], sizeof(double) * n6);
    memcpy(L->mu, keep[3], sizeof(double) * L->n);
  }
  for (int k = 0; k < 4; ++k) free(keep[k]);
  /* the step's preamble (MG:586-601) follows readComputationData (MG:545):
   * it runs once every level has its new geometry, at the next operation */
  if (!rc) c->pending = 1;
  return rc;
}

/* ------------------------------------------------------------------------
 * node-wise routines
 * ------------------------------------------------------------------------ */

/* makePressureField (DEF:13612-13630) */
static void pressure(const or_ctx *C, orl *L, const double *u) {
  double g1 = C->P.gamma - 1.0;
  int n = L->n;
  for (int i = 0; i < n; ++i) {
    double u2 = u[n + i], u3 = u[2 * (size_t)n + i], u4 = u[3 * (size_t)n + i];
    double p = g1 * (u[4 * (size_t)n + i] - 0.5 * (u2 * u2 + u3 * u3 + u4 * u4) / u[i]);
    L->p[i] = p > 0.1 ? p : 0.1;
  }
}

/* limitDensity (DEF:1243-1257) */
static void limit_density(const or_ctx *C, orl *L, double *u1) {
  for (int i = 0; i < L->n; ++i)
    if (!(u1[i] >= C->P.minimum_density)) u1[i] = C->P.minimum_density;
}

/* limitTurbulence (DEF:1259-1290): clamps the argument at zero, then the
 * grid's own u(:,6) at maxTurbulenceValue. */
static void limit_turbulence(const or_ctx *C, orl *L, double *u6) {
  for (int i = 0; i < L->n; ++i) {
    u6[i] = u6[i] > 0.0 ? u6[i] : 0.0;
    if (U(L, 5, i) > C->P.max_turbulence_value) U(L, 5, i) = C->P.max_turbulence_value;
  }
}

/* ------------------------------------------------------------------------
 * boundary conditions
 * ------------------------------------------------------------------------ */

/* setOuterBoundaryConditions (DEF:4018-4251): farfield indicator 3 uses
 * inflowField(:,1), indicator 4 inflowField(:,2). */
static void outer_bc(const or_ctx *C, orl *L, double *u, const double *p) {
  const flite_params *P = &C->P;
  int n = L->n;
  double gm1 = P->gamma - 1.0;
  for (int k = 0; k < 2; ++k) {
    int kstart = k == 0 ? -16 : -14;
    const double *inf = P->inflow_field[k];
    double rho0 = inf[0], U1 = inf[1], U2 = inf[2], U3 = inf[3], E0 = inf[4], p0 = inf[5];
    FOR_RANGE(L, kstart, ib, ip) {
      double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
      double rho = u[ip], m1 = u[n + ip], m2 = u[2 * (size_t)n + ip], m3 = u[3 * (size_t)n + ip];
      double nv = U1 * nx + U2 * ny + U3 * nz;
      if (nv > 0.0) {
        if (P->mach_number > 1.0) {
          u[ip] = rho0;
          u[n + ip] = U1 * rho0;
          u[2 * (size_t)n + ip] = U2 * rho0;
          u[3 * (size_t)n + ip] = U3 * rho0;
          u[4 * (size_t)n + ip] = E0 * rho0;
        } else {
          u[ip] = rho0;
          u[n + ip] = U1 * rho0;
          u[2 * (size_t)n + ip] = U2 * rho0;
          u[3 * (size_t)n + ip] = U3 * rho0;
          u[4 * (size_t)n + ip] = p[ip] / gm1 + 0.5 * (U1 * U1 + U2 * U2 + U3 * U3) * rho0;
        }
        if (P->turbulence_model == 1) u[5 * (size_t)n + ip] = 0.1;
      } else {
        if (P->mach_number < 1.0) u[4 * (size_t)n + ip] = p0 / gm1 + 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
        if (P->turbulence_model == 1) u[5 * (size_t)n + ip] = 0.1;
      }
    }
  }
}

/* setBCsOnSolutionField (DEF:4364-4976), wall / inviscid wall / symmetry
 * branches; engine ranges are empty on the supported meshes. */
static int bc_solution(const or_ctx *C, orl *L, double *u) {
  const flite_params *P = &C->P;
  int n = L->n;
  if (REG(L, -12) <= REG(L, -11) || REG(L, -10) <= REG(L, -9) || REG(L, -8) <= REG(L, -7) ||
      REG(L, -6) <= REG(L, -5))
    return fail(FLITE_E_UNSUPPORTED, "engine boundary nodes are not supported");
  /* coordinateMovement exists on level 1 only (DEF:13882); coarse levels
   * take zero movement */
  const double *cm = L->gnum == 1 ? L->cmv : NULL;
  double dtp = P->physical_time_step;
  if (P->reynolds_number > 0.0) {
    FOR_RANGE(L, -20, ib, ip) {
      /* u(2:4) = u(1) * coordinateMovement / dt (DEF:4408) */
      u[n + ip] = cm ? u[ip] * cm[ip] / dtp : 0.0;
      u[2 * (size_t)n + ip] = cm ? u[ip] * cm[n + ip] / dtp : 0.0;
      u[3 * (size_t)n + ip] = cm ? u[ip] * cm[2 * (size_t)n + ip] / dtp : 0.0;
      if (!P->walls_are_isentropic) u[4 * (size_t)n + ip] = u[ip] * P->inflow_temperature / P->gamma;
      if (P->turbulence_model == 1) u[5 * (size_t)n + ip] = 0.0;
    }
  } else {
    FOR_RANGE(L, -20, ib, ip) {
      double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
      double m1 = u[n + ip], m2 = u[2 * (size_t)n + ip], m3 = u[3 * (size_t)n + ip];
      double nv = m1 * nx + m2 * ny + m3 * nz;
      /* the wall's normal velocity on level 1 of a time-accurate run, times
       * the density of the level's own solution grp%u (DEF:4433-4442) */
      double wn = 0.0;
      if (cm && P->add_time_step)
        wn = (cm[ip] / dtp) * nx + (cm[n + ip] / dtp) * ny + (cm[2 * (size_t)n + ip] / dtp) * nz;
      double m = L->u[ip] * wn;
      u[n + ip] = m1 - nv * nx + m * nx;
      u[2 * (size_t)n + ip] = m2 - nv * ny + m * ny;
      u[3 * (size_t)n + ip] = m3 - nv * nz + m * nz;
      if (P->turbulence_model > 0) u[5 * (size_t)n + ip] = 0.0;
    }
  }
  FOR_RANGE(L, -4, ib, ip) { /* inviscid wall, relaxationParam = 1 */
    double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
    double m1 = u[n + ip], m2 = u[2 * (size_t)n + ip], m3 = u[3 * (size_t)n + ip];
    double nv = m1 * nx + m2 * ny + m3 * nz;
    u[n + ip] = m1 - nv * nx;
    u[2 * (size_t)n + ip] = m2 - nv * ny;
    u[3 * (size_t)n + ip] = m3 - nv * nz;
    if (!P->walls_are_isentropic) {
      double a = u[n + ip], bq = u[2 * (size_t)n + ip], cq = u[3 * (size_t)n + ip];
      u[4 * (size_t)n + ip] = u[ip] * P->inflow_temperature / P->gamma + 0.5 * (a * a + bq * bq + cq * cq) / u[ip];
    }
    if (P->turbulence_model > 0) u[5 * (size_t)n + ip] = 0.0;
  }
  FOR_RANGE(L, -18, ib, ip) { /* symmetry */
    double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
    double m1 = u[n + ip], m2 = u[2 * (size_t)n + ip], m3 = u[3 * (size_t)n + ip];
    double nv = m1 * nx + m2 * ny + m3 * nz;
    u[n + ip] = m1 - nv * nx;
    u[2 * (size_t)n + ip] = m2 - nv * ny;
    u[3 * (size_t)n + ip] = m3 - nv * nz;
  }
  return 0;
}

/* setBCsOnIncrementField (DEF:4978-5140) */
static void bc_increment(const or_ctx *C, orl *L, double *r) {
  const flite_params *P = &C->P;
  int n = L->n;
  if (P->reynolds_number > 0.0) {
    FOR_RANGE(L, -20, ib, ip) {
      r[n + ip] = 0.0;
      r[2 * (size_t)n + ip] = 0.0;
      r[3 * (size_t)n + ip] = 0.0;
      if (!P->walls_are_isentropic) r[4 * (size_t)n + ip] = 0.0;
      if (P->turbulence_model == 1) r[5 * (size_t)n + ip] = 0.0;
    }
  } else {
    FOR_RANGE(L, -20, ib, ip) {
      double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
      double a = r[n + ip], b = r[2 * (size_t)n + ip], c = r[3 * (size_t)n + ip];
      double ni = a * nx + b * ny + c * nz;
      r[n + ip] = a - ni * nx;
      r[2 * (size_t)n + ip] = b - ni * ny;
      r[3 * (size_t)n + ip] = c - ni * nz;
      if (P->turbulence_model == 1) r[5 * (size_t)n + ip] = 0.0;
    }
  }
  FOR_RANGE(L, -4, ib, ip) {
    double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
    double a = r[n + ip], b = r[2 * (size_t)n + ip], c = r[3 * (size_t)n + ip];
    double ni = a * nx + b * ny + c * nz;
    r[n + ip] = a - ni * nx;
    r[2 * (size_t)n + ip] = b - ni * ny;
    r[3 * (size_t)n + ip] = c - ni * nz;
    if (P->turbulence_model == 1) r[5 * (size_t)n + ip] = 0.0;
  }
  FOR_RANGE(L, -18, ib, ip) {
    double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
    double a = r[n + ip], b = r[2 * (size_t)n + ip], c = r[3 * (size_t)n + ip];
    double ni = a * nx + b * ny + c * nz;
    r[n + ip] = a - ni * nx;
    r[2 * (size_t)n + ip] = b - ni * ny;
    r[3 * (size_t)n + ip] = c - ni * nz;
  }
}

/* ------------------------------------------------------------------------
 * edge loops
 * ------------------------------------------------------------------------ */

/* makeTimeSteps (DEF:13335-13610) */
static void timesteps(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  int n = L->n;
  double *dt = L->dt;
  for (int i = 0; i < n; ++i) dt[i] = 0.0;
  for (int i = 0; i < C->nh; ++i) H(C, 0, i) = 0.0;
  double c1 = P->gamma, Re = P->reynolds_number;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double al = wx * wx + wy * wy + wz * wz;
    double ro1 = 1. / U(L, 0, i1), ro2 = 1. / U(L, 0, i2);
    double cc1 = c1 * L->p[i1] * ro1, cc2 = c1 * L->p[i2] * ro2;
    cc1 = cc1 > 0.001 ? cc1 : 0.001;
    cc2 = cc2 > 0.001 ? cc2 : 0.001;
    double vn1 = (wx * U(L, 1, i1) + wy * U(L, 2, i1) + wz * U(L, 3, i1)) * ro1;
    double vn2 = (wx * U(L, 1, i2) + wy * U(L, 2, i2) + wz * U(L, 3, i2)) * ro2;
    dt[i1] = dt[i1] + fabs(vn1) + sqrt(cc1 * al);
    dt[i2] = dt[i2] + fabs(vn2) + sqrt(cc2 * al);
    double visc1, visc2;
    if (P->turbulence_model == 1) {
      visc1 = L->mu[i1] + U(L, 5, i1) / Re;
      visc2 = L->mu[i2] + U(L, 5, i2) / Re;
    } else {
      visc1 = L->mu[i1];
      visc2 = L->mu[i2];
    }
    if (Re > 0.0) {
      double am = 1.0 / Re;
      double vis1 = am > visc1 ? am : visc1, vis2 = am > visc2 ? am : visc2;
      double dv1 = 2.0 * vis1 * al / U(L, 0, i1), dv2 = 2.0 * vis2 * al / U(L, 0, i2);
      H(C, 0, i1) += dv2;
      H(C, 0, i2) += dv1;
    }
  }
  for (int s = 0; s < L->nbs; ++s) {
    int i1 = L->b1[s], i2 = L->b2[s];
    double wx = L->wb[s], wy = L->wb[L->nbs + s], wz = L->wb[2 * (size_t)L->nbs + s];
    double al = wx * wx + wy * wy + wz * wz;
    double ro1 = 1. / U(L, 0, i1), ro2 = 1. / U(L, 0, i2);
    double cc1 = c1 * L->p[i1] * ro1, cc2 = c1 * L->p[i2] * ro2;
    cc1 = cc1 > 0.001 ? cc1 : 0.001;
    cc2 = cc2 > 0.001 ? cc2 : 0.001;
    double vn1 = fabs((wx * U(L, 1, i1) + wy * U(L, 2, i1) + wz * U(L, 3, i1)) * ro1);
    double vn2 = fabs((wx * U(L, 1, i2) + wy * U(L, 2, i2) + wz * U(L, 3, i2)) * ro2);
    double vnm = vn1 + vn2;
    cc1 = sqrt(cc1 * al);
    cc2 = sqrt(cc2 * al);
    double ccm = cc1 + cc2;
    dt[i1] = dt[i1] + vn1 + vnm + cc1 + ccm;
    dt[i2] = dt[i2] + vn2 + vnm + cc2 + ccm;
    double visc1, visc2;
    if (P->turbulence_model == 1) {
      visc1 = L->mu[i1] + U(L, 5, i1) / Re;
      visc2 = L->mu[i2] + U(L, 5, i2) / Re;
    } else {
      visc1 = L->mu[i1];
      visc2 = L->mu[i2];
    }
    if (Re > 0.0) {
      double am = 1.0 / Re;
      double vis1 = am > visc1 ? am : visc1, vis2 = am > visc2 ? am : visc2;
      double dv1 = 2.0 * vis1 * al / U(L, 0, i1), dv2 = 2.0 * vis2 * al / U(L, 0, i2);
      double dvm = dv1 + dv2;
      H(C, 0, i1) = H(C, 0, i1) + dv1 + dvm;
      H(C, 0, i2) = H(C, 0, i2) + dv2 + dvm;
    }
  }
  for (int i = 0; i < n; ++i)
    dt[i] = 0.5 / (dt[i] / L->vol[i] + H(C, 0, i) / (L->vol[i] * L->vol[i]));
}

/* makeArtificialDissipation2 (DEF:11484-11993): zeroes dissipation and the
 * help array; returns early for HLLC on grid 1 or when both factors are
 * zero (DEF:11506-11511).  Otherwise the JST blend of second and fourth
 * differences:
 *   pass 1 (DEF:11531-11563): L_i = sum_j (U_j - U_i) with the energy
 *     difference taken on E + p (help 1-5; help 6 sums the weights, unused);
 *   pass 2 (DEF:11659-11672): sum_j (p_j - p_i) and sum_j (p_i + p_j)
 *     (help 7, 8);
 *   node (DEF:11746-11751): switch nu_i = 12 |help7 / help8| (help 9),
 *     L_i /= nodeConnectivity_i (the number of sides at the node, DEF:14283);
 *   pass 3 (DEF:11788-11939): per side, with dr4 = L_1 - L_2, dr2 = U_1 - U_2,
 *     scalar: d0 = min(V_1/dt_1, V_2/dt_2) / (conn_1 + conn_2),
 *             d2 = d0 k2 max(nu_1, nu_2), d4 = dim(d0 k4, d2);
 *     matrix: dr4 and dr2 replaced by |w| |A| dr4 and |w| |A| dr2 with the
 *             Roe matrix along w (Turkel's form, Harten's entropy fix) and
 *             d2 = k2 max(nu_1, nu_2), d4 = dim(k4, d2);
 *     D_1 += -dr4 d4 + dr2 d2, D_2 += dr4 d4 - dr2 d2.
 * k4 is the coarse-grid factor below grid 1.  Side-length weighting
 * (useDissipationWeighting) is not restated. */
static int dissipation2(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  int n = L->n;
  memset(L->diss, 0, sizeof(double) * 5 * (size_t)n);
  memset(C->help, 0, sizeof(double) * 25 * (size_t)C->nh);
  if (P->flux_type == 4 && L->gnum == 1) return 0;
  if (P->fourth_order_dissipation_factor == 0.0 && P->second_order_dissipation_factor == 0.0) return 0;
  int scheme = L->gnum == 1 ? P->dissipation_scheme : P->coarse_grid_dissipation_scheme;
  if (scheme != 2) return fail(FLITE_E_UNSUPPORTED, "only dissipation scheme 2 (JST) is restated");
  double *conn = zalloc(n);
  for (int e = 0; e < L->ne; ++e) {
    conn[L->e1[e]] += 1.0;
    conn[L->e2[e]] += 1.0;
  }
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double d[5];
    for (int c = 0; c < 5; ++c) d[c] = U(L, c, i1) - U(L, c, i2);
    d[4] = d[4] + L->p[i1] - L->p[i2];
    double df = 1.0;
    for (int c = 0; c < 5; ++c) {
      H(C, c, i1) = H(C, c, i1) - d[c] * df;
      H(C, c, i2) = H(C, c, i2) + d[c] * df;
    }
    H(C, 5, i1) += df;
    H(C, 5, i2) += df;
  }
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double dp = L->p[i1] - L->p[i2], sp = L->p[i1] + L->p[i2];
    H(C, 6, i1) = H(C, 6, i1) - dp;
    H(C, 7, i1) = H(C, 7, i1) + sp;
    H(C, 6, i2) = H(C, 6, i2) + dp;
    H(C, 7, i2) = H(C, 7, i2) + sp;
  }
  for (int i = 0; i < n; ++i) {
    double fact = 1. / conn[i];
    H(C, 8, i) = 12. * fabs(H(C, 6, i) / H(C, 7, i));
    for (int c = 0; c < 5; ++c) H(C, c, i) = fact * H(C, c, i);
  }
  double k2 = P->second_order_dissipation_factor;
  double k4 = L->gnum == 1 ? P->fourth_order_dissipation_factor : P->coarse_grid_dissipation_factor;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double d4v[5], d2v[5], d2, d4;
    for (int c = 0; c < 5; ++c) {
      d4v[c] = H(C, c, i1) - H(C, c, i2);
      d2v[c] = U(L, c, i1) - U(L, c, i2);
    }
    double nu = H(C, 8, i1) > H(C, 8, i2) ? H(C, 8, i1) : H(C, 8, i2);
    if (!P->use_matrix_dissipation) {
      double a = L->vol[i1] / L->dt[i1], b = L->vol[i2] / L->dt[i2];
      double d0 = (a < b ? a : b) / (conn[i1] + conn[i2]);
      d2 = d0 * k2 * nu;
      d4 = d0 * k4;
      d4 = d4 > d2 ? d4 - d2 : 0.0;
    } else {
      double hcf = P->hartens_correction_factor;
      double eps1 = 1. / (hcf > 1.e-5 ? hcf : 1.e-5), gam1 = P->gamma - 1.;
      double rnx = L->w[e], rny = L->w[L->ne + e], rnz = L->w[2 * (size_t)L->ne + e];
      double al = sqrt(rnx * rnx + rny * rny + rnz * rnz), vmo1 = 1. / al;
      rnx *= vmo1;
      rny *= vmo1;
      rnz *= vmo1;
      double rhol = U(L, 0, i1), r1 = 1. / rhol;
      double uxl = U(L, 1, i1) * r1, vyl = U(L, 2, i1) * r1, wzl = U(L, 3, i1) * r1;
      double hl = (U(L, 4, i1) + L->p[i1]) * r1;
      double rhor = U(L, 0, i2), r2 = 1. / rhor;
      double uxr = U(L, 1, i2) * r2, vyr = U(L, 2, i2) * r2, wzr = U(L, 3, i2) * r2;
      double hr = (U(L, 4, i2) + L->p[i2]) * r2;
      double di = sqrt(rhor / rhol), d1 = 1.0 / (di + 1.0);
      double ui = (di * uxr + uxl) * d1, vi = (di * vyr + vyl) * d1, wi = (di * wzr + wzl) * d1;
      double hi = (di * hr + hl) * d1;
      double ci2 = gam1 * (hi - 0.5 * (ui * ui + vi * vi + wi * wi)), ci = sqrt(ci2);
      double af = 0.5 * (ui * ui + vi * vi + wi * wi), ucp = ui * rnx + vi * rny + wi * rnz;
      double l1 = fabs(ucp + ci), l2 = fabs(ucp - ci), l3 = fabs(ucp);
      if (l1 < hcf) l1 = 0.5 * (l1 * l1 * eps1 + hcf);
      if (l2 < hcf) l2 = 0.5 * (l2 * l2 * eps1 + hcf);
      if (l3 < hcf) l3 = 0.5 * (l3 * l3 * eps1 + hcf);
      double s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
      double *dv[2] = {d4v, d2v};
      for (int k = 0; k < 2; ++k) {
        double *q = dv[k];
        double a1 = gam1 * (af * q[0] - ui * q[1] - vi * q[2] - wi * q[3] + q[4]);
        double a2 = -ucp * q[0] + q[1] * rnx + q[2] * rny + q[3] * rnz;
        double cc1 = ((s1 - l3) * a1 / ci2) + (s2 * a2 / ci);
        double cc2 = (s2 * a1 / ci) + (s1 - l3) * a2;
        double t0 = al * (l3 * q[0] + cc1), t1 = al * (l3 * q[1] + cc1 * ui + cc2 * rnx);
        double t2 = al * (l3 * q[2] + cc1 * vi + cc2 * rny), t3 = al * (l3 * q[3] + cc1 * wi + cc2 * rnz);
        double t4 = al * (l3 * q[4] + cc1 * hi + cc2 * ucp);
        q[0] = t0; q[1] = t1; q[2] = t2; q[3] = t3; q[4] = t4;
      }
      d2 = k2 * nu;
      d4 = k4;
      d4 = d4 > d2 ? d4 - d2 : 0.0;
    }
    for (int c = 0; c < 5; ++c) {
      L->diss[(size_t)c * n + i1] = L->diss[(size_t)c * n + i1] - d4v[c] * d4 + d2v[c] * d2;
      L->diss[(size_t)c * n + i2] = L->diss[(size_t)c * n + i2] + d4v[c] * d4 - d2v[c] * d2;
    }
  }
  free(conn);
  return 0;
}

/* makeViscosityTerm, scheme 1 (DEF:5198-5893); result subtracted from
 * rhs(:,2:5) = dissipation. */
static void viscous(or_ctx *C, orl *L, double *rhs) {
  const flite_params *P = &C->P;
  int n = L->n;
  for (int c = 0; c < 12; ++c)
    for (int i = 0; i < n; ++i) H(C, c, i) = 0.0;
  double gog1 = P->gamma / (P->gamma - 1.0);
  double tempConv = (P->gamma - 1.0) * P->mach_number * P->mach_number;
  double Tin = P->inflow_temperature;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double r1 = U(L, 0, i1), u1 = U(L, 1, i1) / r1, v1 = U(L, 2, i1) / r1, w1 = U(L, 3, i1) / r1;
    double t1 = gog1 * L->p[i1] / r1;
    double r2 = U(L, 0, i2), u2 = U(L, 1, i2) / r2, v2 = U(L, 2, i2) / r2, w2 = U(L, 3, i2) / r2;
    double t2 = gog1 * L->p[i2] / r2;
    double s[4] = {u1 + u2, v1 + v2, w1 + w2, t1 + t2};
    double ww[3] = {wx, wy, wz};
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 4; ++q) {
        double f = s[q] * ww[d];
        H(C, 4 * d + q, i1) += f;
        H(C, 4 * d + q, i2) -= f;
      }
  }
  for (int sd = 0; sd < L->nbs; ++sd) {
    int i1 = L->b1[sd], i2 = L->b2[sd];
    double ww[3] = {L->wb[sd], L->wb[L->nbs + sd], L->wb[2 * (size_t)L->nbs + sd]};
    double r1 = U(L, 0, i1), r2 = U(L, 0, i2);
    double a1[4] = {U(L, 1, i1) / r1, U(L, 2, i1) / r1, U(L, 3, i1) / r1, gog1 * L->p[i1] / r1};
    double a2[4] = {U(L, 1, i2) / r2, U(L, 2, i2) / r2, U(L, 3, i2) / r2, gog1 * L->p[i2] / r2};
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 4; ++q) {
        double s = a1[q] + a2[q];
        H(C, 4 * d + q, i1) += (2. * a1[q] + s) * ww[d];
        H(C, 4 * d + q, i2) += (2. * a2[q] + s) * ww[d];
      }
  }
  double oneOverRe = 1. / P->reynolds_number, oneOverPr = 1. / P->prandtl_number;
  double oneOverPrt = 1. / P->turbulent_prandtl_number;
  double twoThirds = 2. / 3., fourThirds = 4. / 3.;
#pragma omp parallel for schedule(static) num_threads(C->threads)
  for (int i = 0; i < n; ++i) {
    double vt = 1. / L->vol[i];
    double dudx = H(C, 0, i) * vt, dvdx = H(C, 1, i) * vt, dwdx = H(C, 2, i) * vt, dTdx = H(C, 3, i) * vt;
    double dudy = H(C, 4, i) * vt, dvdy = H(C, 5, i) * vt, dwdy = H(C, 6, i) * vt, dTdy = H(C, 7, i) * vt;
    double dudz = H(C, 8, i) * vt, dvdz = H(C, 9, i) * vt, dwdz = H(C, 10, i) * vt, dTdz = H(C, 11, i) * vt;
    double T = gog1 * L->p[i] / U(L, 0, i);
    if (T <= 0.001) T = 0.001;
    double mu = oneOverRe * (pow(tempConv * T, 1.5) * ((Tin + 198.6) / (tempConv * Tin * T + 198.6)));
    double k = oneOverPr * mu;
    L->mu[i] = mu;
    L->vort[i] = sqrt((dudy - dvdx) * (dudy - dvdx) + (dvdz - dwdy) * (dvdz - dwdy) + (dwdx - dudz) * (dwdx - dudz));
    L->divg[i] = dudx + dvdy + dwdz;
    if (P->turbulence_model == 1) {
      double tvc = oneOverRe * U(L, 0, i) * U(L, 5, i);
      double ksi = tvc / mu;
      double k3 = ksi * ksi * ksi;
      tvc = tvc * k3 / (k3 + 357.9);
      mu = mu + tvc;
      k = k + tvc * oneOverPrt;
    }
    H(C, 0, i) = mu * (fourThirds * dudx - twoThirds * (dvdy + dwdz));
    H(C, 1, i) = mu * (dudy + dvdx);
    H(C, 2, i) = mu * (dudz + dwdx);
    H(C, 3, i) = mu * (fourThirds * dvdy - twoThirds * (dudx + dwdz));
    H(C, 4, i) = mu * (dvdz + dwdy);
    H(C, 5, i) = mu * (fourThirds * dwdz - twoThirds * (dudx + dvdy));
    H(C, 6, i) = k * dTdx;
    H(C, 7, i) = k * dTdy;
    H(C, 8, i) = k * dTdz;
  }
  if (L->gnum == 1) {
    /* setTemperatureForAdiabaticWall (DEF:5170-5196) */
    FOR_RANGE(L, -20, ib, ip) {
      double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
      double a = H(C, 6, ip), b = H(C, 7, ip), c = H(C, 8, ip);
      double nt = a * nx + b * ny + c * nz;
      H(C, 6, ip) = a - nt * nx;
      H(C, 7, ip) = b - nt * ny;
      H(C, 8, ip) = c - nt * nz;
    }
    /* wallStress (DEF:5622-5638) */
    for (int ib = 0; ib < L->nb; ++ib) {
      int ip = L->nia[ib];
      double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
      L->ws[ip] = H(C, 0, ip) * nx + H(C, 1, ip) * ny + H(C, 2, ip) * nz;
      L->ws[L->nb + ip] = H(C, 1, ip) * nx + H(C, 3, ip) * ny + H(C, 4, ip) * nz;
      L->ws[2 * (size_t)L->nb + ip] = H(C, 2, ip) * nx + H(C, 4, ip) * ny + H(C, 5, ip) * nz;
      L->ws[3 * (size_t)L->nb + ip] = 0.0;
    }
  }
  for (int c = 9; c < 13; ++c)
    for (int i = 0; i < n; ++i) H(C, c, i) = 0.0;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double t11 = H(C, 0, i1), t21 = H(C, 1, i1), t31 = H(C, 2, i1), t41 = H(C, 3, i1), t51 = H(C, 4, i1);
    double t61 = H(C, 5, i1), t71 = H(C, 6, i1), t81 = H(C, 7, i1), t91 = H(C, 8, i1);
    double u11 = U(L, 1, i1) / U(L, 0, i1), u21 = U(L, 2, i1) / U(L, 0, i1), u31 = U(L, 3, i1) / U(L, 0, i1);
    double t12 = H(C, 0, i2), t22 = H(C, 1, i2), t32 = H(C, 2, i2), t42 = H(C, 3, i2), t52 = H(C, 4, i2);
    double t62 = H(C, 5, i2), t72 = H(C, 6, i2), t82 = H(C, 7, i2), t92 = H(C, 8, i2);
    double u12 = U(L, 1, i2) / U(L, 0, i2), u22 = U(L, 2, i2) / U(L, 0, i2), u32 = U(L, 3, i2) / U(L, 0, i2);
    double dt1 = t11 + t12, dt2 = t21 + t22, dt3 = t31 + t32, dt4 = t41 + t42, dt5 = t51 + t52;
    double dt6 = t61 + t62, dt7 = t71 + t72, dt8 = t81 + t82, dt9 = t91 + t92;
    double dut1 = u11 * t11 + u12 * t12 + u21 * t21 + u22 * t22 + u31 * t31 + u32 * t32;
    double dut2 = u11 * t21 + u12 * t22 + u21 * t41 + u22 * t42 + u31 * t51 + u32 * t52;
    double dut3 = u11 * t31 + u12 * t32 + u21 * t51 + u22 * t52 + u31 * t61 + u32 * t62;
    double f2 = wx * dt1 + wy * dt2 + wz * dt3;
    double f3 = wx * dt2 + wy * dt4 + wz * dt5;
    double f4 = wx * dt3 + wy * dt5 + wz * dt6;
    double f5 = wx * (dt7 + dut1) + wy * (dt8 + dut2) + wz * (dt9 + dut3);
    H(C, 9, i1) += f2; H(C, 10, i1) += f3; H(C, 11, i1) += f4; H(C, 12, i1) += f5;
    H(C, 9, i2) -= f2; H(C, 10, i2) -= f3; H(C, 11, i2) -= f4; H(C, 12, i2) -= f5;
    MAG(C, L, 1, i1, f2); MAG(C, L, 2, i1, f3); MAG(C, L, 3, i1, f4); MAG(C, L, 4, i1, f5);
    MAG(C, L, 1, i2, f2); MAG(C, L, 2, i2, f3); MAG(C, L, 3, i2, f4); MAG(C, L, 4, i2, f5);
  }
  for (int sd = 0; sd < L->nbs; ++sd) {
    int ind = L->bi[sd];
    if (ind == 2 || ind == 9) continue;
    int i1 = L->b1[sd], i2 = L->b2[sd];
    double wx = L->wb[sd], wy = L->wb[L->nbs + sd], wz = L->wb[2 * (size_t)L->nbs + sd];
    double t11 = H(C, 0, i1), t21 = H(C, 1, i1), t31 = H(C, 2, i1), t41 = H(C, 3, i1), t51 = H(C, 4, i1);
    double t61 = H(C, 5, i1), t71 = H(C, 6, i1), t81 = H(C, 7, i1), t91 = H(C, 8, i1);
    double u11 = U(L, 1, i1) / U(L, 0, i1), u21 = U(L, 2, i1) / U(L, 0, i1), u31 = U(L, 3, i1) / U(L, 0, i1);
    double t12 = H(C, 0, i2), t22 = H(C, 1, i2), t32 = H(C, 2, i2), t42 = H(C, 3, i2), t52 = H(C, 4, i2);
    double t62 = H(C, 5, i2), t72 = H(C, 6, i2), t82 = H(C, 7, i2), t92 = H(C, 8, i2);
    double u12 = U(L, 1, i2) / U(L, 0, i2), u22 = U(L, 2, i2) / U(L, 0, i2), u32 = U(L, 3, i2) / U(L, 0, i2);
    double dt1 = t11 + t12, dt2 = t21 + t22, dt3 = t31 + t32, dt4 = t41 + t42, dt5 = t51 + t52;
    double dt6 = t61 + t62, dt7 = t71 + t72, dt8 = t81 + t82, dt9 = t91 + t92;
    double dut11 = u11 * t11 + u21 * t21 + u31 * t31, dut12 = u12 * t12 + u22 * t22 + u32 * t32;
    double dut21 = u11 * t21 + u21 * t41 + u31 * t51, dut22 = u12 * t22 + u22 * t42 + u32 * t52;
    double dut31 = u11 * t31 + u21 * t51 + u31 * t61, dut32 = u12 * t32 + u22 * t52 + u32 * t62;
    double dut1 = dut11 + dut12, dut2 = dut21 + dut22, dut3 = dut31 + dut32;
    double f2 = wx * dt1 + wy * dt2 + wz * dt3;
    double f3 = wx * dt2 + wy * dt4 + wz * dt5;
    double f4 = wx * dt3 + wy * dt5 + wz * dt6;
    double f5 = wx * (dt7 + dut1) + wy * (dt8 + dut2) + wz * (dt9 + dut3);
    H(C, 9, i1) = H(C, 9, i1) + f2 + 2.0 * (wx * t11 + wy * t21 + wz * t31);
    H(C, 10, i1) = H(C, 10, i1) + f3 + 2.0 * (wx * t21 + wy * t41 + wz * t51);
    H(C, 11, i1) = H(C, 11, i1) + f4 + 2.0 * (wx * t31 + wy * t51 + wz * t61);
    H(C, 12, i1) = H(C, 12, i1) + f5 + 2.0 * (wx * (t71 + dut11) + wy * (t81 + dut21) + wz * (t91 + dut31));
    H(C, 9, i2) = H(C, 9, i2) + f2 + 2.0 * (wx * t12 + wy * t22 + wz * t32);
    H(C, 10, i2) = H(C, 10, i2) + f3 + 2.0 * (wx * t22 + wy * t42 + wz * t52);
    H(C, 11, i2) = H(C, 11, i2) + f4 + 2.0 * (wx * t32 + wy * t52 + wz * t62);
    H(C, 12, i2) = H(C, 12, i2) + f5 + 2.0 * (wx * (t72 + dut12) + wy * (t82 + dut22) + wz * (t92 + dut32));
    if (C->mag) {
      double m1[4] = {f2 + 2.0 * (wx * t11 + wy * t21 + wz * t31), f3 + 2.0 * (wx * t21 + wy * t41 + wz * t51),
                      f4 + 2.0 * (wx * t31 + wy * t51 + wz * t61),
                      f5 + 2.0 * (wx * (t71 + dut11) + wy * (t81 + dut21) + wz * (t91 + dut31))};
      double m2[4] = {f2 + 2.0 * (wx * t12 + wy * t22 + wz * t32), f3 + 2.0 * (wx * t22 + wy * t42 + wz * t52),
                      f4 + 2.0 * (wx * t32 + wy * t52 + wz * t62),
                      f5 + 2.0 * (wx * (t72 + dut12) + wy * (t82 + dut22) + wz * (t92 + dut32))};
      for (int q = 0; q < 4; ++q) {
        MAG(C, L, 1 + q, i1, m1[q]);
        MAG(C, L, 1 + q, i2, m2[q]);
      }
    }
  }
  for (int c = 1; c < 5; ++c)
    for (int i = 0; i < n; ++i) rhs[(size_t)c * n + i] -= H(C, 8 + c, i);
}

/* computeGradient (DEF:2923-3245): Green-Gauss gradients of (rho, u, v, w,
 * rhoE) into help columns 1-15 (x: 1-5, y: 6-10, z: 11-15). */
static void gradient(or_ctx *C, orl *L) {
  int n = L->n;
  for (int c = 0; c < 15; ++c)
    for (int i = 0; i < n; ++i) H(C, c, i) = 0.0;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double ww[3] = {L->w[e], L->w[L->ne + e], L->w[2 * (size_t)L->ne + e]};
    double r1 = U(L, 0, i1), r2 = U(L, 0, i2);
    double q1[5] = {r1, U(L, 1, i1) / r1, U(L, 2, i1) / r1, U(L, 3, i1) / r1, U(L, 4, i1)};
    double q2[5] = {r2, U(L, 1, i2) / r2, U(L, 2, i2) / r2, U(L, 3, i2) / r2, U(L, 4, i2)};
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 5; ++q) {
        double f = (q1[q] + q2[q]) * ww[d];
        H(C, 5 * d + q, i1) += f;
        H(C, 5 * d + q, i2) -= f;
      }
  }
  for (int s = 0; s < L->nbs; ++s) {
    int i1 = L->b1[s], i2 = L->b2[s];
    double ww[3] = {L->wb[s], L->wb[L->nbs + s], L->wb[2 * (size_t)L->nbs + s]};
    double r1 = U(L, 0, i1), r2 = U(L, 0, i2);
    double q1[5] = {r1, U(L, 1, i1) / r1, U(L, 2, i1) / r1, U(L, 3, i1) / r1, U(L, 4, i1)};
    double q2[5] = {r2, U(L, 1, i2) / r2, U(L, 2, i2) / r2, U(L, 3, i2) / r2, U(L, 4, i2)};
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 5; ++q) {
        double sq = q1[q] + q2[q];
        H(C, 5 * d + q, i1) += (2. * q1[q] + sq) * ww[d];
        H(C, 5 * d + q, i2) += (2. * q2[q] + sq) * ww[d];
      }
  }
  for (int i = 0; i < n; ++i) {
    double vt = 1. / L->vol[i];
    for (int c = 0; c < 15; ++c) H(C, c, i) *= vt;
  }
}

/* LocalMinMax (DEF:3248-3459): help 16-20 = min over edges - q_i,
 * help 21-25 = max over edges - q_i (kept as min of negatives). */
static void local_minmax(or_ctx *C, orl *L) {
  int n = L->n;
  for (int c = 15; c < 25; ++c)
    for (int i = 0; i < n; ++i) H(C, c, i) = 1.0e15;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double r1 = U(L, 0, i1), r2 = U(L, 0, i2);
    double q1[5] = {r1, U(L, 1, i1) / r1, U(L, 2, i1) / r1, U(L, 3, i1) / r1, U(L, 4, i1)};
    double q2[5] = {r2, U(L, 1, i2) / r2, U(L, 2, i2) / r2, U(L, 3, i2) / r2, U(L, 4, i2)};
    for (int q = 0; q < 5; ++q) {
      double mn = q1[q] < q2[q] ? q1[q] : q2[q];
      double mx = -q1[q] < -q2[q] ? -q1[q] : -q2[q];
      if (mn < H(C, 15 + q, i1)) H(C, 15 + q, i1) = mn;
      if (mn < H(C, 15 + q, i2)) H(C, 15 + q, i2) = mn;
      if (mx < H(C, 20 + q, i1)) H(C, 20 + q, i1) = mx;
      if (mx < H(C, 20 + q, i2)) H(C, 20 + q, i2) = mx;
    }
  }
  for (int i = 0; i < n; ++i) {
    double r = U(L, 0, i);
    double q[5] = {r, U(L, 1, i) / r, U(L, 2, i) / r, U(L, 3, i) / r, U(L, 4, i)};
    for (int c = 0; c < 5; ++c) {
      H(C, 15 + c, i) = H(C, 15 + c, i) - q[c];
      H(C, 20 + c, i) = -H(C, 20 + c, i) - q[c];
    }
  }
}

/* LED local clipping inside makeRHS (DEF:1852-1889) */
static double led(double a, double bmin, double bmax, int nsig) {
  double c = a, a1 = 2.0 * a;
  if (a1 > 0.0) {
    double inv = 0.0;
    if (bmax > 0.0) inv = 1.0 / bmax;
    double t = a1 * inv, tn = t * t;
    if (nsig == 4) tn = tn * tn;
    c = 0.5 * bmax * (t) / pow(1. + tn, 1.0 / nsig);
  } else if (a1 < 0.0) {
    double inv = 0.0;
    if (bmin < 0.0) inv = 1.0 / bmin;
    double t = a1 * inv, tn = t * t;
    if (nsig == 4) tn = tn * tn;
    c = 0.5 * bmin * (t) / pow(1. + tn, 1.0 / nsig);
  }
  return c;
}

/* HLLC interior flux of makeRHS (DEF:1794-2088) for one side; st1/st2 are
 * (rho, rhou, rhov, rhow, rhoE) with pressures pr1/pr2; x1/x2 coordinates;
 * g1/g2 pointers to help rows are read through the callback indices. */
static void hllc_side(const flite_params *P, int order, int autoclip, int nsig,
                      const double *s1, double pr1, const double *s2, double pr2,
                      double wx, double wy, double wz, const double *dx3,
                      const double *gr1, const double *gr2, double *f, double *usel) {
  double gam = P->gamma;
  double r1 = s1[0], rinv1 = 1. / s1[0], ru1 = s1[1], rv1 = s1[2], rw1 = s1[3];
  double u1 = rinv1 * ru1, v1 = rinv1 * rv1, w1 = rinv1 * rw1, E1 = s1[4], rh1 = s1[4] + pr1;
  double r2 = s2[0], rinv2 = 1. / s2[0], ru2 = s2[1], rv2 = s2[2], rw2 = s2[3];
  double u2 = rinv2 * ru2, v2 = rinv2 * rv2, w2 = rinv2 * rw2, E2 = s2[4], rh2 = s2[4] + pr2;
  /* central part, kept for the branch-free fallthrough (DEF:1798-1802) */
  f[0] = wx * (ru1 + ru2) + wy * (rv1 + rv2) + wz * (rw1 + rw2);
  f[1] = wx * (ru1 * u1 + pr1 + ru2 * u2 + pr2) + wy * (ru1 * v1 + ru2 * v2) + wz * (ru1 * w1 + ru2 * w2);
  f[2] = wx * (rv1 * u1 + rv2 * u2) + wy * (rv1 * v1 + pr1 + rv2 * v2 + pr2) + wz * (rv1 * w1 + rv2 * w2);
  f[3] = wx * (rw1 * u1 + rw2 * u2) + wy * (rw1 * v1 + rw2 * v2) + wz * (rw1 * w1 + pr1 + rw2 * w2 + pr2);
  f[4] = wx * (rh1 * u1 + rh2 * u2) + wy * (rh1 * v1 + rh2 * v2) + wz * (rh1 * w1 + rh2 * w2);
  double roe1 = r1, roe2 = r2, re1 = E1, re2 = E2;
  wx = 2.0 * wx;
  wy = 2.0 * wy;
  wz = 2.0 * wz;
  double rnorm = sqrt(wx * wx + wy * wy + wz * wz);
  double rnorm1 = rnorm;
  if (rnorm > 0.0) rnorm = 1.0 / rnorm;
  double rnx = wx * rnorm, rny = wy * rnorm, rnz = wz * rnorm;
  if (order > 1) {
    double dx = dx3[0], dy = dx3[1], dz = dx3[2];
    double hsize = 0.5;
    double hlength = 0.5 * P->hllc_clipping * sqrt(dx * dx + dy * dy + dz * dz);
    double dpvl[5];
    for (int j = 0; j < 5; ++j) {
      double dpv1 = hsize * (gr1[j] * dx + gr1[j + 5] * dy + gr1[j + 10] * dz);
      double dpv11 = hsize * (gr2[j] * dx + gr2[j + 5] * dy + gr2[j + 10] * dz);
      if (autoclip == 1) {
        dpv1 = led(dpv1, gr1[j + 15], gr1[j + 20], nsig);
        dpv11 = -led(-dpv11, gr2[j + 15], gr2[j + 20], nsig);
      }
      double a = dpv1, b = dpv11, c;
      if (a * b <= 0.0) {
        c = 0.0;
      } else {
        c = fabs(a) <= fabs(b) ? a : b;
        if (fabs(c) > hlength && hlength > 0.0) c = hlength * c / fabs(c);
      }
      dpvl[j] = c;
    }
    if (roe1 + dpvl[0] > 0.1) roe1 = roe1 + dpvl[0];
    u1 = u1 + dpvl[1];
    v1 = v1 + dpvl[2];
    w1 = w1 + dpvl[3];
    if (re1 + dpvl[4] > 0.0001) re1 = re1 + dpvl[4];
    pr1 = (gam - 1) * (re1 - 0.5 * (u1 * u1 + v1 * v1 + w1 * w1) * roe1);
    pr1 = pr1 > 0.1 ? pr1 : 0.1;
    rh1 = re1 + pr1;
    rinv1 = 1 / roe1;
    ru1 = roe1 * u1;
    rv1 = roe1 * v1;
    rw1 = roe1 * w1;
    if (roe2 - dpvl[0] > 0.1) roe2 = roe2 - dpvl[0];
    u2 = u2 - dpvl[1];
    v2 = v2 - dpvl[2];
    w2 = w2 - dpvl[3];
    if (re2 - dpvl[4] > 0.0001) re2 = re2 - dpvl[4];
    pr2 = (gam - 1) * (re2 - 0.5 * (u2 * u2 + v2 * v2 + w2 * w2) * roe2);
    pr2 = pr2 > 0.1 ? pr2 : 0.1;
    rh2 = re2 + pr2;
    rinv2 = 1 / roe2;
    ru2 = roe2 * u2;
    rv2 = roe2 * v2;
    rw2 = roe2 * w2;
  }
  double nv1 = u1 * rnx + v1 * rny + w1 * rnz;
  double nv2 = u2 * rnx + v2 * rny + w2 * rnz;
  double beta0 = sqrt((gam - 1) / (2 * gam));
  double rroe = sqrt(roe2 / roe1);
  double ut = (u1 + u2 * rroe) / (1 + rroe);
  double vt = (v1 + v2 * rroe) / (1 + rroe);
  double wt = (w1 + w2 * rroe) / (1 + rroe);
  double qt = ut * rnx + vt * rny + wt * rnz;
  double ht = (rh1 * rinv1 + rh2 * rinv2 * rroe) / (1 + rroe);
  double ct = (gam - 1) * (ht - 0.5 * (ut * ut + vt * vt + wt * wt));
  ct = sqrt(fabs(ct));
  double c1 = sqrt((gam - 1) * fabs(rh1 * rinv1 - 0.5 * (u1 * u1 + v1 * v1 + w1 * w1)));
  double c2 = sqrt((gam - 1) * fabs(rh2 * rinv2 - 0.5 * (u2 * u2 + v2 * v2 + w2 * w2)));
  double sl = fmin(nv1 - beta0 * c1, qt - ct);
  double sr = fmax(nv2 + beta0 * c2, qt + ct);
  sl = fmin(sl, nv2 - beta0 * c2);
  sr = fmax(sr, nv1 + beta0 * c1);
  double sm = roe2 * nv2 * (sr - nv2) - roe1 * nv1 * (sl - nv1);
  sm = sm + pr1 - pr2;
  double fact = roe2 * (sr - nv2) - roe1 * (sl - nv1);
  if (fact != 0.0) fact = 1.0 / fact;
  sm = fact * sm;
  double fl[5], fr[5];
  fl[0] = rnx * (ru1) + rny * (rv1) + rnz * (rw1);
  fl[1] = rnx * (ru1 * u1 + pr1) + rny * (ru1 * v1) + rnz * (ru1 * w1);
  fl[2] = rnx * (rv1 * u1) + rny * (rv1 * v1 + pr1) + rnz * (rv1 * w1);
  fl[3] = rnx * (rw1 * u1) + rny * (rw1 * v1) + rnz * (rw1 * w1 + pr1);
  fl[4] = rnx * (rh1 * u1) + rny * (rh1 * v1) + rnz * (rh1 * w1);
  fr[0] = rnx * (ru2) + rny * (rv2) + rnz * (rw2);
  fr[1] = rnx * (ru2 * u2 + pr2) + rny * (ru2 * v2) + rnz * (ru2 * w2);
  fr[2] = rnx * (rv2 * u2) + rny * (rv2 * v2 + pr2) + rnz * (rv2 * w2);
  fr[3] = rnx * (rw2 * u2) + rny * (rw2 * v2) + rnz * (rw2 * w2 + pr2);
  fr[4] = rnx * (rh2 * u2) + rny * (rh2 * v2) + rnz * (rh2 * w2);
  double plstar = roe1 * (nv1 - sl) * (nv1 - sm) + pr1;
  double prstar = roe2 * (nv2 - sr) * (nv2 - sm) + pr2;
  double ls[5], rs[5];
  double factl = sl - sm;
  if (factl != 0.0) {
    factl = 1.0 / factl;
    ls[0] = roe1 * (sl - nv1) * factl;
    ls[1] = ((sl - nv1) * roe1 * u1 + (plstar - pr1) * rnx) * factl;
    ls[2] = ((sl - nv1) * roe1 * v1 + (plstar - pr1) * rny) * factl;
    ls[3] = ((sl - nv1) * roe1 * w1 + (plstar - pr1) * rnz) * factl;
    ls[4] = ((sl - nv1) * re1 - pr1 * nv1 + plstar * sm) * factl;
  } else {
    ls[0] = roe1; ls[1] = ru1; ls[2] = rv1; ls[3] = rw1; ls[4] = re1;
  }
  double factr = sr - sm;
  if (factr != 0.0) {
    factr = 1.0 / factr;
    rs[0] = roe2 * (sr - nv2) * factr;
    rs[1] = ((sr - nv2) * roe2 * u2 + (prstar - pr2) * rnx) * factr;
    rs[2] = ((sr - nv2) * roe2 * v2 + (prstar - pr2) * rny) * factr;
    rs[3] = ((sr - nv2) * roe2 * w2 + (prstar - pr2) * rnz) * factr;
    rs[4] = ((sr - nv2) * re2 - pr2 * nv2 + prstar * sm) * factr;
  } else {
    rs[0] = roe2; rs[1] = ru2; rs[2] = rv2; rs[3] = rw2; rs[4] = re2;
  }
  double st1[5] = {roe1, ru1, rv1, rw1, re1}, st2[5] = {roe2, ru2, rv2, rw2, re2};
  /* the state each branch leaves in r1, ru1, rv1, rw1, E1 for the GCL term
   * (DEF:2033-2073): density and energy of the node, momentum as
   * reconstructed, or the star state */
  double g1[5] = {r1, ru1, rv1, rw1, E1}, g2[5] = {r2, ru2, rv2, rw2, E2};
  const double *gs = g1;
  if (sl > 0.0) {
    for (int k = 0; k < 5; ++k) f[k] = fl[k] * rnorm1;
  } else if (sl <= 0.0 && sm > 0.0) {
    for (int k = 0; k < 5; ++k) f[k] = (fl[k] + sl * (ls[k] - st1[k])) * rnorm1;
    gs = ls;
  } else if (sm <= 0.0 && sr >= 0.0) {
    for (int k = 0; k < 5; ++k) f[k] = (fr[k] + sr * (rs[k] - st2[k])) * rnorm1;
    gs = rs;
  } else if (sr < 0.0) {
    for (int k = 0; k < 5; ++k) f[k] = fr[k] * rnorm1;
    gs = g2;
  }
  if (usel)
    for (int k = 0; k < 5; ++k) usel[k] = gs[k];
}

int or_hllc_flux(const flite_params *p, const double ul[5], double pl,
                 const double ur[5], double pr, const double w[3], double f[5]) {
  double z[3] = {0, 0, 0};
  hllc_side(p, 1, 0, 2, ul, pl, ur, pr, w[0], w[1], w[2], z, NULL, NULL, f, NULL);
  return 0;
}

int or_hllc_side(const flite_params *p, int order, int autoclip, int nsig,
                 const double ul[5], double pl, const double ur[5], double pr,
                 const double w[3], const double dx[3], const double g1[25],
                 const double g2[25], double f[5]) {
  if (order < 1 || order > 2 || (nsig != 2 && nsig != 4)) return fail(FLITE_E_ARG, "bad order or nsig");
  hllc_side(p, order, autoclip, nsig, ul, pl, ur, pr, w[0], w[1], w[2], dx, g1, g2, f, NULL);
  return 0;
}

double or_led(double a, double bmin, double bmax, int nsig) { return led(a, bmin, bmax, nsig); }

/* farfield Roe flux of one side end against the inflow state
 * (DEF:2258-2446) */
static void farfield_roe(const flite_params *P, const double *inf, const double *s,
                         double pr, double rnx, double rny, double rnz, double *f) {
  double gam1 = P->gamma - 1.0;
  double ro = inf[0], ro1 = 1. / ro, uo = inf[1], vo = inf[2], wo = inf[3], eo = inf[4], po = inf[5];
  double rho = s[0], rinv = 1. / rho;
  double rh = s[4] + pr;
  double uu = s[1] * rinv, vv = s[2] * rinv, ww = s[3] * rinv;
  double eps = s[4], h = rh * rinv;
  double nv = uu * rnx + vv * rny + ww * rnz;
  double fl1 = rho * nv, fl2 = pr * rnx + rho * uu * nv, fl3 = pr * rny + rho * vv * nv;
  double fl4 = pr * rnz + rho * ww * nv, fl5 = rh * nv;
  double rhor = ro, epsr = eo * ro, presr = po, hr = eo + po * ro1;
  double unr = uo * rnx + vo * rny + wo * rnz;
  double fr1 = rhor * unr, fr2 = rnx * presr + rhor * uo * unr, fr3 = rny * presr + rhor * vo * unr;
  double fr4 = rnz * presr + rhor * wo * unr, fr5 = (epsr + presr) * unr;
  double di = sqrt(rhor / rho), d1 = 1.0 / (di + 1.0);
  double ui = (di * uo + uu) * d1, vi = (di * vo + vv) * d1, wi = (di * wo + ww) * d1, hi = (di * hr + h) * d1;
  double ci2 = gam1 * (hi - 0.5 * (ui * ui + vi * vi + wi * wi));
  ci2 = ci2 > 1.0e-5 ? ci2 : 1.0e-5;
  double ci = sqrt(ci2), af = 0.5 * (ui * ui + vi * vi + wi * wi), ucp = ui * rnx + vi * rny + wi * rnz;
  double du1 = rhor - rho, du2 = rhor * uo - rho * uu, du3 = rhor * vo - rho * vv;
  double du4 = rhor * wo - rho * ww, du5 = epsr - eps;
  double rl1 = fabs(ucp + ci), rl2 = fabs(ucp - ci), rl3 = fabs(ucp);
  /* epslm = 0: the entropy-fix branches never fire */
  double s1 = 0.5 * (rl1 + rl2), s2 = 0.5 * (rl1 - rl2);
  double al1x = gam1 * (af * du1 - ui * du2 - vi * du3 - wi * du4 + du5);
  double al2x = -ucp * du1 + du2 * rnx + du3 * rny + du4 * rnz;
  double cc1 = ((s1 - rl3) * al1x / ci2) + (s2 * al2x / ci);
  double cc2 = (s2 * al1x / ci) + (s1 - rl3) * al2x;
  f[0] = 0.5 * (fr1 + fl1) - 0.5 * (rl3 * du1 + cc1);
  f[1] = 0.5 * (fr2 + fl2) - 0.5 * (rl3 * du2 + cc1 * ui + cc2 * rnx);
  f[2] = 0.5 * (fr3 + fl3) - 0.5 * (rl3 * du3 + cc1 * vi + cc2 * rny);
  f[3] = 0.5 * (fr4 + fl4) - 0.5 * (rl3 * du4 + cc1 * wi + cc2 * rnz);
  f[4] = 0.5 * (fr5 + fl5) - 0.5 * (rl3 * du5 + cc1 * hi + cc2 * ucp);
}

int or_farfield_flux(const flite_params *p, int indicator, const double s[5], double pr,
                     const double n[3], double f[5]) {
  if (indicator != 3 && indicator != 4) return fail(FLITE_E_ARG, "farfield indicator must be 3 or 4");
  farfield_roe(p, p->inflow_field[indicator - 3], s, pr, n[0], n[1], n[2], f);
  return 0;
}

/* makeRHS (DEF:1525-2590), HLLC interior flux + boundary loop; rhs must be
 * zeroed by the caller (the reference does grp%rhs = 0 before the call). */
/* Interior side flux of fluxType 1-3 (makeRHS, DEF:1664-1792), no
 * reconstruction (the JST dissipation supplies the higher-order damping):
 *   1 central (DEF:1664-1669): w . (F(U_1) + F(U_2));
 *   2 local Lax-Friedrichs (DEF:1670-1724) and 3 Roe (DEF:1725-1792), both
 *     |w| (F_1 + F_2 + D) along w/|w|.  The reference reads rho1/rho2 before
 *     any assignment in these branches (DEF:1683,1694; 1744,1752), reads the
 *     Roe enthalpies h1/h2 unassigned (DEF:1763), and subtracts the upwind
 *     term D (lambda (U_1 - U_2), or the Roe matrix times U_1 - U_2), which
 *     makes it anti-dissipative.  Restated patched: rho = u(:,1),
 *     h = (E + p)/rho and D added, which is the textbook flux.  Roe's
 *     entropy fix is inactive (epslm = 0, DEF:1580).  Sides with
 *     |w| <= 1e-10 keep no flux (the reference keeps the previous side's). */
static void simple_side(const flite_params *P, int ft, const double *s1, double pr1, const double *s2, double pr2,
                        double wx, double wy, double wz, double *f) {
  double r1 = s1[0], rinv1 = 1. / s1[0], ru1 = s1[1], rv1 = s1[2], rw1 = s1[3], E1 = s1[4];
  double u1 = rinv1 * ru1, v1 = rinv1 * rv1, w1 = rinv1 * rw1, rh1 = E1 + pr1;
  double r2 = s2[0], rinv2 = 1. / s2[0], ru2 = s2[1], rv2 = s2[2], rw2 = s2[3], E2 = s2[4];
  double u2 = rinv2 * ru2, v2 = rinv2 * rv2, w2 = rinv2 * rw2, rh2 = E2 + pr2;
  if (ft == 1) {
    f[0] = wx * (ru1 + ru2) + wy * (rv1 + rv2) + wz * (rw1 + rw2);
    f[1] = wx * (ru1 * u1 + pr1 + ru2 * u2 + pr2) + wy * (ru1 * v1 + ru2 * v2) + wz * (ru1 * w1 + ru2 * w2);
    f[2] = wx * (rv1 * u1 + rv2 * u2) + wy * (rv1 * v1 + pr1 + rv2 * v2 + pr2) + wz * (rv1 * w1 + rv2 * w2);
    f[3] = wx * (rw1 * u1 + rw2 * u2) + wy * (rw1 * v1 + rw2 * v2) + wz * (rw1 * w1 + pr1 + rw2 * w2 + pr2);
    f[4] = wx * (rh1 * u1 + rh2 * u2) + wy * (rh1 * v1 + rh2 * v2) + wz * (rh1 * w1 + rh2 * w2);
    return;
  }
  for (int c = 0; c < 5; ++c) f[c] = 0.0;
  double rnorm = sqrt(wx * wx + wy * wy + wz * wz);
  if (!(rnorm > 1.0e-10)) return;
  double rnx = wx / rnorm, rny = wy / rnorm, rnz = wz / rnorm;
  double nv1 = u1 * rnx + v1 * rny + w1 * rnz, nv2 = u2 * rnx + v2 * rny + w2 * rnz;
  double F1[5] = {r1 * nv1, pr1 * rnx + ru1 * nv1, pr1 * rny + rv1 * nv1, pr1 * rnz + rw1 * nv1, rh1 * nv1};
  double F2[5] = {r2 * nv2, pr2 * rnx + ru2 * nv2, pr2 * rny + rv2 * nv2, pr2 * rnz + rw2 * nv2, rh2 * nv2};
  double du[5] = {r1 - r2, ru1 - ru2, rv1 - rv2, rw1 - rw2, E1 - E2};
  double D[5];
  if (ft == 2) {
    double c1 = sqrt(P->gamma * pr1 * rinv1), c2 = sqrt(P->gamma * pr2 * rinv2);
    double l1 = fabs(nv1 + c1), l12 = fabs(nv2 + c2);
    if (l1 < l12) l1 = l12;
    double l2 = fabs(nv1 - c1), l22 = fabs(nv2 - c2);
    if (l2 < l22) l2 = l22;
    double lam = l1;
    if (lam < l2) lam = l2;
    for (int c = 0; c < 5; ++c) D[c] = lam * du[c];
  } else {
    double gam1 = P->gamma - 1.0;
    double h1 = rh1 * rinv1, h2 = rh2 * rinv2;
    double di = sqrt(r1 / r2), d1 = 1.0 / (di + 1.0);
    double ui = (di * u1 + u2) * d1, vi = (di * v1 + v2) * d1, wi = (di * w1 + w2) * d1;
    double hi = (di * h1 + h2) * d1;
    double ci2 = gam1 * (hi - 0.5 * (ui * ui + vi * vi + wi * wi));
    ci2 = ci2 > 1.0e-5 ? ci2 : 1.0e-5;
    double ci = sqrt(ci2), af = 0.5 * (ui * ui + vi * vi + wi * wi), ucp = ui * rnx + vi * rny + wi * rnz;
    double l1 = fabs(ucp + ci), l2 = fabs(ucp - ci), l3 = fabs(ucp);
    double s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
    double a1 = gam1 * (af * du[0] - ui * du[1] - vi * du[2] - wi * du[3] + du[4]);
    double a2 = -ucp * du[0] + du[1] * rnx + du[2] * rny + du[3] * rnz;
    double cc1 = ((s1 - l3) * a1 / ci2) + (s2 * a2 / ci);
    double cc2 = (s2 * a1 / ci) + (s1 - l3) * a2;
    D[0] = l3 * du[0] + cc1;
    D[1] = l3 * du[1] + cc1 * ui + cc2 * rnx;
    D[2] = l3 * du[2] + cc1 * vi + cc2 * rny;
    D[3] = l3 * du[3] + cc1 * wi + cc2 * rnz;
    D[4] = l3 * du[4] + cc1 * hi + cc2 * ucp;
  }
  for (int c = 0; c < 5; ++c) f[c] = rnorm * ((F1[c] + F2[c]) + D[c]);
}

int or_simple_flux(const flite_params *p, int flux_type, const double ul[5], double pl,
                   const double ur[5], double pr, const double w[3], double f[5]) {
  if (!p || flux_type < 1 || flux_type > 3) return fail(FLITE_E_ARG, "flux_type must be 1-3");
  simple_side(p, flux_type, ul, pl, ur, pr, w[0], w[1], w[2], f);
  return 0;
}

static int make_rhs(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  int n = L->n;
  if (P->flux_type < 1 || P->flux_type > 4) return fail(FLITE_E_UNSUPPORTED, "fluxType must be 1-4");
  int order = P->flux_type == 4 ? P->hllc_flux_order : 1, autoclip = P->automatic_clipping;
  int nsig = P->reynolds_number < 1 ? 4 : 2;
  if (order > 1) {
    if (P->gradient_method != 1) return fail(FLITE_E_UNSUPPORTED, "least-squares gradients are not restated");
    gradient(C, L);
    if (autoclip == 1) local_minmax(C, L);
  }
  /* the side fluxes first (independent: threads when or_set_threads asked
   * for them), then their scatter in side order, so every thread count
   * gives the same bits */
  double *fe = (double *)malloc(sizeof(double) * 5 * (size_t)(L->ne > 0 ? L->ne : 1));
  if (!fe) return fail(FLITE_E_ARG, "out of host memory");
#pragma omp parallel for schedule(static) num_threads(C->threads)
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double gr1[25], gr2[25];
    double s1[5], s2[5];
    for (int c = 0; c < 5; ++c) { s1[c] = U(L, c, i1); s2[c] = U(L, c, i2); }
    double dx[3];
    for (int c = 0; c < 3; ++c) dx[c] = L->xyz[(size_t)c * n + i2] - L->xyz[(size_t)c * n + i1];
    if (order > 1)
      for (int c = 0; c < 25; ++c) { gr1[c] = H(C, c, i1); gr2[c] = H(C, c, i2); }
    double f[5], usel[5];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    /* wgcl = GCLWeightsArray / physicalTimestep (DEF:1662) */
    double wgcl = L->gcl ? L->gcl[e] / P->physical_time_step : 0.0;
    if (P->flux_type == 4) {
      hllc_side(P, order, autoclip, nsig, s1, L->p[i1], s2, L->p[i2], wx, wy, wz, dx, gr1, gr2, f, usel);
      if (wgcl != 0.0) /* the state the HLLC branch left behind (DEF:2081-2085) */
        for (int c = 0; c < 5; ++c) f[c] -= usel[c] * wgcl;
    } else {
      simple_side(P, P->flux_type, s1, L->p[i1], s2, L->p[i2], wx, wy, wz, f);
      /* -(U1 + U2) wgcl: always for the central flux, inside the |w| > 1e-10
       * branch for LLF and Roe (DEF:1665-1669, 1718-1722, 1787-1791) */
      if (wgcl != 0.0 && (P->flux_type == 1 || sqrt(wx * wx + wy * wy + wz * wz) > 1.0e-10))
        for (int c = 0; c < 5; ++c) f[c] -= (s1[c] + s2[c]) * wgcl;
    }
    for (int c = 0; c < 5; ++c) fe[5 * (size_t)e + c] = f[c];
  }
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    const double *f = fe + 5 * (size_t)e;
    for (int c = 0; c < 5; ++c) {
      RHS(L, c, i1) += f[c];
      RHS(L, c, i2) -= f[c];
      MAG(C, L, c, i1, f[c]);
      MAG(C, L, c, i2, f[c]);
    }
  }
  free(fe);
  for (int s = 0; s < L->nbs; ++s) {
    int i1 = L->b1[s], i2 = L->b2[s], ind = L->bi[s];
    double wx = L->wb[s], wy = L->wb[L->nbs + s], wz = L->wb[2 * (size_t)L->nbs + s];
    double s1[5], s2[5];
    for (int c = 0; c < 5; ++c) { s1[c] = U(L, c, i1); s2[c] = U(L, c, i2); }
    if (L->gclb) { /* GCL terms of every boundary side (DEF:2124-2179) */
      double g1 = L->gclb[s] / P->physical_time_step, g2 = L->gclb[L->nbs + s] / P->physical_time_step;
      for (int c = 0; c < 5; ++c) {
        if (P->boundary_term == 1) {
          RHS(L, c, i1) -= (3. * s1[c] + s2[c]) * g1;
          RHS(L, c, i2) -= (s1[c] + 3. * s2[c]) * g2;
          MAG(C, L, c, i1, (3. * s1[c] + s2[c]) * g1);
          MAG(C, L, c, i2, (s1[c] + 3. * s2[c]) * g2);
        } else {
          RHS(L, c, i1) -= s1[c] * (4.0 * g1);
          RHS(L, c, i2) -= s2[c] * (4.0 * g2);
        }
      }
    }
    double f1[5], f2[5];
    double rnorm = sqrt(wx * wx + wy * wy + wz * wz);
    if (ind <= 2 || ind == 9) {
      if (!(rnorm > 1.0e-10)) continue;
      double rnx = wx / rnorm, rny = wy / rnorm, rnz = wz / rnorm;
      double pr1 = L->p[i1], rho1 = s1[0], rinv1 = 1. / rho1, rh1 = s1[4] + pr1;
      double u1 = s1[1] * rinv1, v1 = s1[2] * rinv1, w1 = s1[3] * rinv1;
      double nv1 = u1 * rnx + v1 * rny + w1 * rnz;
      f1[0] = rho1 * nv1;
      f1[1] = pr1 * rnx + rho1 * u1 * nv1;
      f1[2] = pr1 * rny + rho1 * v1 * nv1;
      f1[3] = pr1 * rnz + rho1 * w1 * nv1;
      f1[4] = rh1 * nv1;
      double pr2 = L->p[i2], rho2 = s2[0], rinv2 = 1. / rho2, rh2 = s2[4] + pr2;
      double u2 = s2[1] * rinv2, v2 = s2[2] * rinv2, w2 = s2[3] * rinv2;
      double nv2 = u2 * rnx + v2 * rny + w2 * rnz;
      f2[0] = rho2 * nv2;
      f2[1] = pr2 * rnx + rho2 * u2 * nv2;
      f2[2] = pr2 * rny + rho2 * v2 * nv2;
      f2[3] = pr2 * rnz + rho2 * w2 * nv2;
      f2[4] = rh2 * nv2;
    } else if (ind <= 4) {
      if (!(rnorm > 1.0e-10)) continue;
      double rnx = wx / rnorm, rny = wy / rnorm, rnz = wz / rnorm;
      farfield_roe(P, P->inflow_field[ind - 3], s1, L->p[i1], rnx, rny, rnz, f1);
      farfield_roe(P, P->inflow_field[ind - 3], s2, L->p[i2], rnx, rny, rnz, f2);
    } else {
      return fail(FLITE_E_UNSUPPORTED, "engine boundary sides are not restated");
    }
    for (int c = 0; c < 5; ++c) {
      if (P->boundary_term == 1) {
        RHS(L, c, i1) += rnorm * (3. * f1[c] + f2[c]);
        RHS(L, c, i2) += rnorm * (f1[c] + 3. * f2[c]);
        MAG(C, L, c, i1, rnorm * 3. * f1[c]);
        MAG(C, L, c, i1, rnorm * f2[c]);
        MAG(C, L, c, i2, rnorm * f1[c]);
        MAG(C, L, c, i2, rnorm * 3. * f2[c]);
      } else {
        RHS(L, c, i1) += rnorm * 4. * f1[c];
        RHS(L, c, i2) += rnorm * 4. * f2[c];
      }
    }
  }
  return 0;
}

/* makeSARHS (DEF:6809-6950): upwinded convection of nu~ into rhs(:,6) */
static void sa_rhs(or_ctx *C, orl *L) {
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double rinv1 = 1. / U(L, 0, i1), t1 = U(L, 5, i1);
    double u1 = rinv1 * U(L, 1, i1), v1 = rinv1 * U(L, 2, i1), w1 = rinv1 * U(L, 3, i1);
    double rinv2 = 1. / U(L, 0, i2), t2 = U(L, 5, i2);
    double u2 = rinv2 * U(L, 1, i2), v2 = rinv2 * U(L, 2, i2), w2 = rinv2 * U(L, 3, i2);
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double fc = wx * (u1 * t1 + u2 * t2) + wy * (v1 * t1 + v2 * t2) + wz * (w1 * t1 + w2 * t2);
    if (fabs(t1 - t2) > 1.0e-20) {
      double ax = (u2 * t2 - u1 * t1) / (t1 - t2);
      double ay = (v2 * t2 - v1 * t1) / (t1 - t2);
      double az = (w2 * t2 - w1 * t1) / (t1 - t2);
      fc = fc + fabs(wx * ax + wy * ay + wz * az) * (t1 - t2);
    }
    RHS(L, 5, i1) += fc;
    RHS(L, 5, i2) -= fc;
    MAG(C, L, 5, i1, fc);
    MAG(C, L, 5, i2, fc);
  }
}

/* makeSASourceTerm (DEF:9285-9370) */
static void sa_source(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  double Re = P->reynolds_number, oneOverRe = 1. / Re;
  double cb1 = 0.1355, cb2 = 0.622, kappa = 0.41, k2 = kappa * kappa, sigma = 2. / 3.;
  double cw1 = cb1 / k2 + (1.0 + cb2) / sigma, cw2 = 0.3, cw3 = 2.0, cv1 = 7.1;
  double ct3 = 1.1, ct4 = 2.0;
  double cv13 = cv1 * cv1 * cv1;
  double cw36 = pow(cw3, 6);
  double oneOverSix = 1. / 6.;
  double cFact = exp(oneOverSix * log(1.0 + cw36));
  for (int i = 0; i < L->n; ++i) {
    double vt = L->vol[i];
    double mc = U(L, 5, i);
    mc = mc > 0.0 ? mc : 0.0;
    if (L->mu[i] < 0.1 / Re) L->mu[i] = 0.1 * Re;
    double ksi = U(L, 0, i) * oneOverRe * mc / L->mu[i];
    ksi = ksi > 0.001 ? ksi : 0.001;
    double ft2 = ct3 * exp(-ct4 * ksi * ksi);
    double k3 = ksi * ksi * ksi;
    double d = L->wd[i];
    d = d > 1.0e-10 ? d : 1.0e-10;
    double muOverd = mc / d, d2 = d * d;
    double fv1 = k3 / (k3 + cv13);
    double fv2 = 1.0 - (ksi / (1.0 + ksi * fv1));
    double Sc = 1.0 * L->vort[i] + oneOverRe * mc * fv2 / (k2 * d2);
    double r;
    if (fabs(Sc) > 1.0e-10) {
      r = oneOverRe * mc / (Sc * k2 * d2);
      if (r < 0.0) r = 0.0;
      if (r > 100.0) r = 100.0;
    } else {
      r = 100.0;
    }
    double r6 = r * r * r * r * r * r;
    double g = r + cw2 * (r6 - r);
    double g6 = g * g * g * g * g * g;
    double rad = (1.0 + cw36) / (g6 + cw36);
    double fw = rad > 1.0e-10 ? g * exp(oneOverSix * log(rad)) : cFact;
    RHS(L, 5, i) = RHS(L, 5, i) - vt * cb1 * (1.0 - ft2) * Sc * mc;
    RHS(L, 5, i) = RHS(L, 5, i) + oneOverRe * vt * (cw1 * fw - cb1 * ft2 / k2) * muOverd * muOverd;
    RHS(L, 5, i) = RHS(L, 5, i) - vt * mc * L->divg[i];
    MAG(C, L, 5, i, vt * cb1 * (1.0 - ft2) * Sc * mc);
    MAG(C, L, 5, i, oneOverRe * vt * (cw1 * fw - cb1 * ft2 / k2) * muOverd * muOverd);
    MAG(C, L, 5, i, vt * mc * L->divg[i]);
  }
}

/* makeSADiffusion (DEF:7436-7808) */
static void sa_diffusion(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  int n = L->n;
  double sigma = 2. / 3., cb2 = 0.622, Re = P->reynolds_number, oneOverRe = 1. / Re;
  for (int i = 0; i < n; ++i) L->tdiff[i] = 0.0;
  for (int c = 8; c < 14; ++c)
    for (int i = 0; i < n; ++i) H(C, c, i) = 0.0;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double s = U(L, 5, i1) + U(L, 5, i2);
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    H(C, 8, i1) += wx * s; H(C, 8, i2) -= wx * s;
    H(C, 9, i1) += wy * s; H(C, 9, i2) -= wy * s;
    H(C, 10, i1) += wz * s; H(C, 10, i2) -= wz * s;
  }
  for (int sd = 0; sd < L->nbs; ++sd) {
    int i1 = L->b1[sd], i2 = L->b2[sd];
    double ww[3] = {L->wb[sd], L->wb[L->nbs + sd], L->wb[2 * (size_t)L->nbs + sd]};
    double t1 = U(L, 5, i1), t2 = U(L, 5, i2);
    for (int d = 0; d < 3; ++d) {
      if (P->boundary_term == 1) {
        H(C, 8 + d, i1) += (3. * t1 + t2) * ww[d];
        H(C, 8 + d, i2) += (t1 + 3. * t2) * ww[d];
      } else {
        H(C, 8 + d, i1) += 4. * t1 * ww[d];
        H(C, 8 + d, i2) += 4. * t2 * ww[d];
      }
    }
  }
  for (int i = 0; i < C->nh; ++i) H(C, 0, i) = 0.0;
  double fact = oneOverRe * cb2 / sigma;
  for (int i = 0; i < n; ++i) {
    double vt = 1. / L->vol[i];
    double dx = H(C, 8, i) * vt, dy = H(C, 9, i) * vt, dz = H(C, 10, i) * vt;
    H(C, 11, i) = dx;
    H(C, 12, i) = dy;
    H(C, 13, i) = dz;
    double vp = H(C, 8, i) * dx + H(C, 9, i) * dy + H(C, 10, i) * dz;
    H(C, 0, i) = H(C, 0, i) - fact * vp;
  }
  fact = oneOverRe / sigma;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double v1 = Re * L->mu[i1] / U(L, 0, i1) + U(L, 5, i1);
    double v2 = Re * L->mu[i2] / U(L, 0, i2) + U(L, 5, i2);
    double fd1 = v1 * (wx * H(C, 11, i1) + wy * H(C, 12, i1) + wz * H(C, 13, i1));
    double fd2 = v2 * (wx * H(C, 11, i2) + wy * H(C, 12, i2) + wz * H(C, 13, i2));
    L->tdiff[i1] = L->tdiff[i1] - fact * (fd1 + fd2);
    L->tdiff[i2] = L->tdiff[i2] + fact * (fd1 + fd2);
    MAG(C, L, 5, i1, fact * (fd1 + fd2));
    MAG(C, L, 5, i2, fact * (fd1 + fd2));
  }
  for (int sd = 0; sd < L->nbs; ++sd) {
    int i1 = L->b1[sd], i2 = L->b2[sd];
    double wx = L->wb[sd], wy = L->wb[L->nbs + sd], wz = L->wb[2 * (size_t)L->nbs + sd];
    double v1 = Re * L->mu[i1] / U(L, 0, i1) + U(L, 5, i1);
    double v2 = Re * L->mu[i2] / U(L, 0, i2) + U(L, 5, i2);
    double fd1 = v1 * (wx * H(C, 11, i1) + wy * H(C, 12, i1) + wz * H(C, 13, i1));
    double fd2 = v2 * (wx * H(C, 11, i2) + wy * H(C, 12, i2) + wz * H(C, 13, i2));
    if (P->boundary_term == 1) {
      L->tdiff[i1] = L->tdiff[i1] - fact * (3. * fd1 + fd2);
      L->tdiff[i2] = L->tdiff[i2] - fact * (fd1 + 3. * fd2);
      MAG(C, L, 5, i1, fact * (3. * fd1 + fd2));
      MAG(C, L, 5, i2, fact * (fd1 + 3. * fd2));
    } else {
      L->tdiff[i1] = L->tdiff[i1] - fact * 4. * fd1;
      L->tdiff[i2] = L->tdiff[i2] - fact * 4. * fd2;
    }
  }
  for (int i = 0; i < n; ++i) {
    L->tdiff[i] = L->tdiff[i] + H(C, 0, i);
    MAG(C, L, 5, i, H(C, 0, i));
  }
}

/* makeTurbulentTimeSteps (DEF:12878-13099) */
static void sa_timesteps(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  int n = L->n;
  double oneOverSigma = 3.0 / 2.0, oneOverK2 = 1.0 / (0.41 * 0.41), cb1 = 0.1355, cb2 = 0.622, ct3 = 1.1;
  double cw1 = cb1 * oneOverK2 + (1.0 + cb2) * oneOverSigma;
  double oneOverRe = 1.0 / P->reynolds_number;
  double *dt = L->dtt;
  for (int i = 0; i < n; ++i) dt[i] = 0.0;
  for (int i = 0; i < C->nh; ++i) H(C, 0, i) = 0.0;
  for (int e = 0; e < L->ne; ++e) {
    int i1 = L->e1[e], i2 = L->e2[e];
    double wx = L->w[e], wy = L->w[L->ne + e], wz = L->w[2 * (size_t)L->ne + e];
    double al = wx * wx + wy * wy + wz * wz;
    double ro1 = 1. / U(L, 0, i1), ro2 = 1. / U(L, 0, i2);
    double vn1 = (wx * U(L, 1, i1) + wy * U(L, 2, i1) + wz * U(L, 3, i1)) * ro1;
    double vn2 = (wx * U(L, 1, i2) + wy * U(L, 2, i2) + wz * U(L, 3, i2)) * ro2;
    vn1 = P->turb_k * sqrt(al) + fabs(vn1);
    vn2 = P->turb_k * sqrt(al) + fabs(vn2);
    dt[i1] += vn1;
    dt[i2] += vn2;
    double vis1 = oneOverSigma * (L->mu[i1] + 3.0 * cb2 * U(L, 5, i1) * oneOverRe);
    double vis2 = oneOverSigma * (L->mu[i2] + 3.0 * cb2 * U(L, 5, i2) * oneOverRe);
    H(C, 0, i1) += 2.0 * vis2 * al;
    H(C, 0, i2) += 2.0 * vis1 * al;
  }
  for (int s = 0; s < L->nbs; ++s) {
    int i1 = L->b1[s], i2 = L->b2[s];
    double wx = L->wb[s], wy = L->wb[L->nbs + s], wz = L->wb[2 * (size_t)L->nbs + s];
    double al = wx * wx + wy * wy + wz * wz;
    double ro1 = 1. / U(L, 0, i1), ro2 = 1. / U(L, 0, i2);
    double vn1 = (wx * U(L, 1, i1) + wy * U(L, 2, i1) + wz * U(L, 3, i1)) * ro1;
    double vn2 = (wx * U(L, 1, i2) + wy * U(L, 2, i2) + wz * U(L, 3, i2)) * ro2;
    vn1 = P->turb_k * sqrt(al) + fabs(vn1);
    vn2 = P->turb_k * sqrt(al) + fabs(vn2);
    double vnm = vn1 + vn2;
    dt[i1] = dt[i1] + vn1 + vnm;
    dt[i2] = dt[i2] + vn2 + vnm;
    double vis1 = oneOverSigma * (L->mu[i1] + 3.0 * cb2 * U(L, 5, i1) * oneOverRe);
    double vis2 = oneOverSigma * (L->mu[i2] + 3.0 * cb2 * U(L, 5, i2) * oneOverRe);
    double dv1 = 2.0 * vis1 * al, dv2 = 2.0 * vis2 * al, dvm = dv1 + dv2;
    H(C, 0, i1) = H(C, 0, i1) + dv1 + dvm;
    H(C, 0, i2) = H(C, 0, i2) + dv2 + dvm;
  }
  for (int i = 0; i < n; ++i) {
    dt[i] = 0.5 / (dt[i] / L->vol[i] + H(C, 0, i) / (L->vol[i] * L->vol[i]));
    double d = dt[i];
    double di = L->wd[i] > 1.0e-10 ? 1.0 / L->wd[i] : 1.0e10;
    di = di * di;
    double S = fabs(L->divg[i]) + cb1 * (L->vort[i] + (1.0 + ct3) * U(L, 5, i) * oneOverK2 * di * oneOverRe) +
               cw1 * U(L, 5, i) * di * oneOverRe;
    dt[i] = d / (1.0 + 0.5 * S * d);
  }
}

/* ------------------------------------------------------------------------
 * smoother
 * ------------------------------------------------------------------------ */

/* solveSystem + solveTurbulenceSystem (DEF:1407-1522), steady flow */
/* addTimeDifferences (DEF:1184-1241): the dual-time source of a fixed
 * mesh (previous volumes = volumes), BDF2 on grid 1 with timeScheme 2,
 * BDF1 otherwise; components 1:5, 1:6 with Spalart-Allmaras */
static void time_differences(or_ctx *C, orl *L) {
  const flite_params *P = &C->P;
  if (!P->add_time_step || P->flow_type == 0) return;
  int n = L->n, nc = P->turbulence_model == 1 ? 6 : 5;
  double dtp = P->physical_time_step;
  for (int i = 0; i < n; ++i) {
    double V = L->vol[i];
    /* nodeVolumeP, nodeVolumeP2: the volumes at the last two physical times
     * (equal to V on a fixed mesh) */
    double VP = L->volp ? L->volp[i] : V, VP2 = L->volp2 ? L->volp2[i] : V;
    for (int c = 0; c < nc; ++c) {
      size_t k = (size_t)c * n + i;
      if (L->gnum == 1 && P->time_scheme == 2)
        L->rhs[k] = L->rhs[k] + (1.5 * V * L->u[k] - 2.0 * VP * L->uold[k] + 0.5 * VP2 * L->uold2[k]) / dtp;
      else
        L->rhs[k] = L->rhs[k] + V * (L->u[k] - L->uold[k]) / dtp;
    }
  }
}

/* solveSystem (DEF:1407-1479); unsteady: dt <- dt / (1 + 1.5 dt / dtp)
 * for the mean flow (DEF:1427-1436), the SA time step unchanged */
static void solve_system(or_ctx *C, orl *L, double *r) {
  int n = L->n;
  double tdc = C->P.flow_type != 0 ? 1.5 : 0.0;
  for (int i = 0; i < n; ++i) {
    double dt = L->dt[i];
    if (C->P.flow_type != 0) dt = dt / (1.0 + tdc * dt / C->P.physical_time_step);
    for (int c = 0; c < 5; ++c) r[(size_t)c * n + i] = dt * r[(size_t)c * n + i] / L->vol[i];
  }
  if (C->P.turbulence_model == 1)
    for (int i = 0; i < n; ++i) r[5 * (size_t)n + i] = L->dtt[i] * r[5 * (size_t)n + i] / L->vol[i];
}

/* updateSolutionVector (DEF:1292-1405); u may alias delu */
static void update(or_ctx *C, orl *L, double *u, const double *uprev, double *delu, int rk) {
  const flite_params *P = &C->P;
  int n = L->n;
  double dtJ = rk > 0 ? P->cfl_number * P->rk_coefficients[rk - 1] : P->cfl_number;
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < n; ++i) delu[(size_t)c * n + i] = dtJ * delu[(size_t)c * n + i];
  if (P->turbulence_model > 0) {
    double dtt = rk > 0 ? P->turbulent_cfl_number * P->rk_coefficients[rk - 1] : P->turbulent_cfl_number;
    if (P->turbulence_model == 1)
      for (int i = 0; i < n; ++i) delu[5 * (size_t)n + i] = dtt * delu[5 * (size_t)n + i];
  }
  for (int i = 0; i < n; ++i) {
    double inc = delu[i];
    if (fabs(inc / uprev[i]) < P->max_increment_factor) {
      u[i] = uprev[i] - inc;
    } else {
      double sg = inc / fabs(inc);
      u[i] = uprev[i] - sg * P->max_increment_factor * uprev[i];
    }
    for (int c = 1; c < 5; ++c) u[(size_t)c * n + i] = uprev[(size_t)c * n + i] - delu[(size_t)c * n + i];
    if (P->turbulence_model == 1) {
      double t = uprev[5 * (size_t)n + i] - delu[5 * (size_t)n + i];
      u[5 * (size_t)n + i] = t > 0.0 ? t : 0.0;
    }
  }
  for (int i = 0; i < n; ++i)
    if (U(L, 0, i) < P->minimum_density) U(L, 0, i) = P->minimum_density;
}

/* singleIteration (DEF:1002-1181); the reference's 3 stages with alpha =
 * (0.6, 0.6, 1.0) are flite_params' defaults, other stage counts (up to 5)
 * run the same stage with the given factors */
static int single_iteration(or_ctx *C, orl *L, int calc_diss, int calc_turb, int iteration) {
  const flite_params *P = &C->P;
  int n = L->n, rc;
  int sa = P->turbulence_model == 1 && calc_turb;
  (void)iteration; /* triggerTurbulenceField: no trip nodes on the supported grids */
  if (calc_diss) {
    pressure(C, L, L->u);
    for (int i = 0; i < n; ++i) L->p[i] = fabs(L->p[i]);
    timesteps(C, L);
    if ((rc = dissipation2(C, L))) return rc;
    if (P->reynolds_number > 0.0) {
      viscous(C, L, L->diss);
      if (sa) {
        sa_timesteps(C, L);
        sa_diffusion(C, L);
      }
    }
  }
  for (int k = 1; k <= P->rk_stages; ++k) {
    pressure(C, L, L->u);
    memset(L->rhs, 0, sizeof(double) * 6 * (size_t)n);
    if ((rc = make_rhs(C, L))) return rc;
    time_differences(C, L);
    if (sa) {
      sa_rhs(C, L);
      sa_source(C, L);
    }
    for (int c = 0; c < 5; ++c)
      for (int i = 0; i < n; ++i) RHS(L, c, i) += L->diss[(size_t)c * n + i];
    if (sa)
      for (int i = 0; i < n; ++i) RHS(L, 5, i) += L->tdiff[i];
    solve_system(C, L, L->rhs);
    bc_increment(C, L, L->rhs);
    update(C, L, L->u, L->uprev, L->rhs, k);
    if ((rc = bc_solution(C, L, L->u))) return rc;
    if (L->gnum > 1) {
      double a = P->rk_coefficients[k - 1];
      for (int c = 0; c < 5; ++c)
        for (int i = 0; i < n; ++i) U(L, c, i) = U(L, c, i) - a * L->src[(size_t)c * n + i];
      if (sa)
        for (int i = 0; i < n; ++i) U(L, 5, i) = U(L, 5, i) - a * L->src[5 * (size_t)n + i];
    }
    limit_density(C, L, L->u);
    if (calc_turb && P->turbulence_model == 1) limit_turbulence(C, L, L->u + 5 * (size_t)n);
    outer_bc(C, L, L->u, L->p);
  }
  return 0;
}

int or_do_time_iterations(or_ctx *c, int level, int steps, int calc_diss, int calc_turb, int iteration) {
  settle(c);
  if (!c || level < 1 || level > c->nl || !c->L[level - 1].uploaded) return fail(FLITE_E_ARG, "bad level");
  orl *L = &c->L[level - 1];
  int n = L->n;
  for (int s = 0; s < steps; ++s) {
    memcpy(L->uprev, L->u, sizeof(double) * 5 * (size_t)n);
    if (c->P.turbulence_model == 1) memcpy(L->uprev + 5 * (size_t)n, L->u + 5 * (size_t)n, sizeof(double) * n);
    int rc = single_iteration(c, L, calc_diss, calc_turb, iteration);
    if (rc) return rc;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * multigrid transfer
 * ------------------------------------------------------------------------ */

/* mapFineToCoarse (DEF:12572-12690) over ncomp columns */
static void map_f2c(const orl *F, const orl *Cl, const double *fu, double *cu, int ncomp) {
  for (int c = 0; c < ncomp; ++c)
    for (int i = 0; i < Cl->n; ++i) cu[(size_t)c * Cl->n + i] = 0.0;
  for (int i = 0; i < Cl->nfine; ++i) {
    int k = Cl->f2c[i];
    for (int c = 0; c < ncomp; ++c) cu[(size_t)c * Cl->n + k] += Cl->rw[i] * fu[(size_t)c * F->n + i];
  }
}

/* mapCoarseToFine, injection (DEF:12692-12790) */
static void map_c2f(const orl *Cl, const orl *F, const double *cu, double *fu, int ncomp) {
  for (int i = 0; i < Cl->nfine; ++i) {
    int k = Cl->f2c[i];
    for (int c = 0; c < ncomp; ++c) fu[(size_t)c * F->n + i] = cu[(size_t)c * Cl->n + k];
  }
}

/* restrictGrid (DEF:12141-12430) */
int or_restrict(or_ctx *C, int fine_level, int calc_turb) {
  settle(C);
  if (!C || fine_level < 1 || fine_level >= C->nl) return fail(FLITE_E_ARG, "bad level");
  orl *F = &C->L[fine_level - 1], *K = &C->L[fine_level];
  const flite_params *P = &C->P;
  int nf = F->n, nc = K->n, rc;
  int sa = P->turbulence_model == 1;
  int sac = sa && calc_turb;
  /* coarse: L_H I_h^H u_h */
  map_f2c(F, K, F->u, K->u, 6);
  map_f2c(F, K, F->mu, K->mu, 1);
  pressure(C, K, K->u);
  if ((rc = bc_solution(C, K, K->u))) return rc;
  outer_bc(C, K, K->u, K->p);
  memcpy(K->uprev, K->u, sizeof(double) * 5 * (size_t)nc);
  if (sa) memcpy(K->uprev + 5 * (size_t)nc, K->u + 5 * (size_t)nc, sizeof(double) * nc);
  pressure(C, K, K->u);
  for (int i = 0; i < nc; ++i) K->p[i] = fabs(K->p[i]);
  timesteps(C, K);
  if ((rc = dissipation2(C, K))) return rc;
  if (P->reynolds_number > 0.0) viscous(C, K, K->diss);
  memset(K->rhs, 0, sizeof(double) * 6 * (size_t)nc);
  if ((rc = make_rhs(C, K))) return rc;
  time_differences(C, K); /* DEF:12208, on the restricted state */
  if (sac) {
    sa_timesteps(C, K);
    sa_diffusion(C, K);
    sa_rhs(C, K);
    sa_source(C, K);
  }
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nc; ++i) RHS(K, c, i) += K->diss[(size_t)c * nc + i];
  if (sa)
    for (int i = 0; i < nc; ++i) RHS(K, 5, i) += K->tdiff[i];
  solve_system(C, K, K->rhs);
  bc_increment(C, K, K->rhs);
  update(C, K, K->u, K->uprev, K->rhs, 0);
  if ((rc = bc_solution(C, K, K->u))) return rc;
  outer_bc(C, K, K->u, K->p);
  for (size_t q = 0; q < 6 * (size_t)nc; ++q) K->src[q] = K->u[q] - K->uprev[q];
  memcpy(K->u, K->uprev, sizeof(double) * 6 * (size_t)nc);
  /* fine: I_h^H L_h u_h as an increment */
  pressure(C, F, F->u);
  for (int i = 0; i < nf; ++i) F->p[i] = fabs(F->p[i]);
  if ((rc = dissipation2(C, F))) return rc;
  if (P->reynolds_number > 0.0) viscous(C, F, F->diss);
  memset(F->rhs, 0, sizeof(double) * 6 * (size_t)nf);
  if ((rc = make_rhs(C, F))) return rc;
  time_differences(C, F); /* DEF:12285 */
  if (sac) {
    sa_diffusion(C, F);
    sa_rhs(C, F);
    sa_source(C, F);
  }
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nf; ++i) RHS(F, c, i) += F->diss[(size_t)c * nf + i];
  if (sa)
    for (int i = 0; i < nf; ++i) RHS(F, 5, i) += F->tdiff[i];
  solve_system(C, F, F->rhs);
  bc_increment(C, F, F->rhs);
  update(C, F, F->rhs, F->uprev, F->rhs, 0);
  if ((rc = bc_solution(C, F, F->rhs))) return rc;
  if (F->gnum > 1) {
    for (int c = 0; c < 5; ++c)
      for (int i = 0; i < nf; ++i) RHS(F, c, i) -= F->src[(size_t)c * nf + i];
    if (sa)
      for (int i = 0; i < nf; ++i) RHS(F, 5, i) -= F->src[5 * (size_t)nf + i];
  }
  if (sa) limit_turbulence(C, F, F->rhs + 5 * (size_t)nf);
  outer_bc(C, F, F->rhs, F->p);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nf; ++i) RHS(F, c, i) -= UP(F, c, i);
  if (sa)
    for (int i = 0; i < nf; ++i) RHS(F, 5, i) -= UP(F, 5, i);
  map_f2c(F, K, F->rhs, K->rhs, 6);
  bc_increment(C, K, K->rhs);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nc; ++i) K->src[(size_t)c * nc + i] -= RHS(K, c, i);
  if (sa)
    for (int i = 0; i < nc; ++i) K->src[5 * (size_t)nc + i] -= RHS(K, 5, i);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nc; ++i) U(K, c, i) += RHS(K, c, i);
  if (sa)
    for (int i = 0; i < nc; ++i) U(K, 5, i) += RHS(K, 5, i);
  return 0;
}

/* prolongateGrid (DEF:12432-12570) */
int or_prolongate(or_ctx *C, int coarse_level) {
  settle(C);
  if (!C || coarse_level < 2 || coarse_level > C->nl) return fail(FLITE_E_ARG, "bad level");
  orl *K = &C->L[coarse_level - 1], *F = &C->L[coarse_level - 2];
  const flite_params *P = &C->P;
  int nc = K->n, nf = F->n, rc;
  int sa = P->turbulence_model == 1;
  map_f2c(F, K, F->u, K->uprev, 6);
  if ((rc = bc_solution(C, K, K->uprev))) return rc;
  outer_bc(C, K, K->uprev, K->p);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nc; ++i) RHS(K, c, i) = U(K, c, i) - UP(K, c, i);
  if (sa)
    for (int i = 0; i < nc; ++i) RHS(K, 5, i) = U(K, 5, i) - UP(K, 5, i);
  bc_increment(C, K, K->rhs);
  map_c2f(K, F, K->rhs, F->rhs, 6);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < nf; ++i) RHS(F, c, i) = P->prolongation_relaxation * RHS(F, c, i);
  if (sa)
    for (int i = 0; i < nf; ++i) RHS(F, 5, i) = P->turb_prolongation_relaxation * RHS(F, 5, i);
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < F->nb; ++i) RHS(F, c, i) = P->multigrid_bc_relaxation * RHS(F, c, i);
  int ncomp = (F->gnum < P->number_of_turbulence_grids && sa) ? 6 : 5;
  for (int i = 0; i < nf; ++i)
    for (int c = 0; c < ncomp; ++c) U(F, c, i) = U(F, c, i) + RHS(F, c, i);
  if ((rc = bc_solution(C, F, F->u))) return rc;
  outer_bc(C, F, F->u, F->p);
  for (int i = 0; i < nf; ++i) {
    if (sa) U(F, 5, i) = U(F, 5, i) > 0.0 ? U(F, 5, i) : 0.0;
    U(F, 0, i) = U(F, 0, i) > P->minimum_density ? U(F, 0, i) : P->minimum_density;
  }
  return 0;
}

int or_map_u_fine_to_coarse(or_ctx *C, int fine_level) {
  settle(C);
  if (!C || fine_level < 1 || fine_level >= C->nl) return fail(FLITE_E_ARG, "bad level");
  map_f2c(&C->L[fine_level - 1], &C->L[fine_level], C->L[fine_level - 1].u, C->L[fine_level].u, 6);
  return 0;
}

/* oneCycle (MG:907-1060), schemes <= 0, 1 and 3 */
int or_cycle(or_ctx *C, int cycle) {
  settle(C);
  const flite_params *P = &C->P;
  int ng = C->nl, rc;
  int ct = cycle <= P->number_of_turbulence_steps;
  int nrs = P->number_of_relaxation_steps;
  if (P->multigrid_scheme <= 0) {
    for (int i = 1; i < ng; ++i) {
      orl *F = &C->L[i - 1];
      memcpy(F->uprev, F->u, sizeof(double) * 6 * (size_t)F->n);
      map_f2c(F, &C->L[i], F->u, C->L[i].u, 6);
    }
    memset(C->L[ng - 1].src, 0, sizeof(double) * 6 * (size_t)C->L[ng - 1].n);
    if ((rc = or_do_time_iterations(C, ng, nrs, 1, ct, cycle))) return rc;
    for (int i = ng; i >= 2; --i) map_c2f(&C->L[i - 1], &C->L[i - 2], C->L[i - 1].u, C->L[i - 2].u, 6);
  } else if (P->multigrid_scheme == 1) {
    for (int i = 1; i < ng; ++i) {
      if ((rc = or_do_time_iterations(C, i, nrs, i == 1, ct, cycle))) return rc;
      if ((rc = or_restrict(C, i, ct))) return rc;
      or_map_u_fine_to_coarse(C, i);
    }
    if ((rc = or_do_time_iterations(C, ng, nrs, ng == 1, ct, cycle))) return rc;
    for (int i = ng; i >= 2; --i) {
      if ((rc = or_prolongate(C, i))) return rc;
      if ((rc = or_do_time_iterations(C, i - 1, nrs, (i - 1) == 1, ct, cycle))) return rc;
    }
  } else if (P->multigrid_scheme == 3) {
    int sweeps = ng == 1 ? 1 : ng - 1;
    int finest = 1;
    for (int s = 1; s <= sweeps; ++s) {
      for (int j = finest; j <= ng - 1; ++j) {
        if ((rc = or_do_time_iterations(C, j, nrs, j == finest, ct, cycle))) return rc;
        if ((rc = or_restrict(C, j, ct))) return rc;
      }
      if ((rc = or_do_time_iterations(C, ng, nrs, ng == 1, ct, cycle))) return rc;
      if (finest == 1)
        finest = ng >= 2 ? ng - 1 : ng;
      else
        finest = finest - 1;
      for (int j = ng; j >= finest + 1; --j)
        if ((rc = or_prolongate(C, j))) return rc;
    }
  } else {
    return fail(FLITE_E_UNSUPPORTED, "multigrid scheme not restated");
  }
  return 0;
}

/* multiGridCycles preamble (MG:586-602): laminar viscosity, and u and
 * uold restricted to every coarse level */
static void preamble(or_ctx *C) {
  double mu0 = C->P.reynolds_number > 0.0 ? 1. / C->P.reynolds_number : 0.0;
  for (int l = 0; l < C->nl; ++l) {
    orl *L = &C->L[l];
    for (int i = 0; i < L->n; ++i) L->mu[i] = mu0;
    if (l > 0) {
      map_f2c(&C->L[l - 1], L, C->L[l - 1].u, L->u, 6);
      map_f2c(&C->L[l - 1], L, C->L[l - 1].uold, L->uold, 6);
    }
  }
}

/* the preamble a mesh replacement left due (or_level_geometry) */
static void settle(or_ctx *C) {
  if (C && C->pending) {
    C->pending = 0;
    preamble(C);
  }
}

/* setUpInitialField2 freestream start (DEF:3940-3996): uold and uold2 take
 * the freestream before the boundary conditions act on u (DEF:3980-3994) */
int or_init_freestream(or_ctx *C) {
  C->pending = 0;
  orl *L = &C->L[0];
  const flite_params *P = &C->P;
  int n = L->n, rc;
  for (int c = 0; c < 5; ++c)
    for (int i = 0; i < n; ++i) U(L, c, i) = P->inflow_field[0][c];
  if (P->turbulence_model == 1)
    for (int i = 0; i < n; ++i) U(L, 5, i) = 0.1;
  memcpy(L->uold, L->u, sizeof(double) * 6 * (size_t)n);
  memcpy(L->uold2, L->u, sizeof(double) * 6 * (size_t)n);
  pressure(C, L, L->u);
  pressure(C, L, L->u);
  if ((rc = bc_solution(C, L, L->u))) return rc;
  preamble(C);
  return 0;
}

/* a start from a given level-1 state (restart, DEF:3866-3873: uold and
 * uold2 take it too) followed by the preamble */
int or_init_from_state(or_ctx *C) {
  C->pending = 0;
  orl *L = &C->L[0];
  memcpy(L->uold, L->u, sizeof(double) * 6 * (size_t)L->n);
  memcpy(L->uold2, L->u, sizeof(double) * 6 * (size_t)L->n);
  preamble(C);
  return 0;
}

/* end of a physical time step (MG:889-890): grid 1's uold2 <- uold <- u;
 * then the preamble of the next step (MG:586-601, inside the physical-step
 * loop): laminar viscosity reset, u and uold restricted to every coarse
 * level */
int or_advance_physical_time(or_ctx *C) {
  settle(C);
  if (!C || !C->L[0].uploaded) return fail(FLITE_E_STATE, "level 1 not uploaded");
  orl *L = &C->L[0];
  memcpy(L->uold2, L->uold, sizeof(double) * 6 * (size_t)L->n);
  memcpy(L->uold, L->u, sizeof(double) * 6 * (size_t)L->n);
  preamble(C);
  return 0;
}

/* writeResiduals reductions (MG:1271-1445) + calculateLiftAndDrag
 * (DEF:14710-14806) on grid 1 */
int or_monitor(or_ctx *C, double res[7], double maxres[7], int maxidx[7], double forces[9]) {
  settle(C);
  orl *L = &C->L[0];
  const flite_params *P = &C->P;
  int n = L->n;
  int nres = P->turbulence_model == 1 ? 6 : 5;
  for (int j = 0; j < 7; ++j) { res[j] = 0; maxres[j] = 0; maxidx[j] = 0; }
  for (int j = 0; j < nres; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      double d = U(L, j, i) - UP(L, j, i);
      s += d * d;
    }
    res[j] = sqrt(s);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nres; ++j) {
      double d = fabs(U(L, j, i) - UP(L, j, i));
      if (d > maxres[j]) { maxres[j] = d; maxidx[j] = i; }
    }
  double PI = 4.0 * atan(1.0);
  double ar = P->alpha * PI / 180., br = P->beta * PI / 180.;
  double ft[3] = {cos(ar) * cos(br), sin(br), sin(ar)};
  double fn1[3] = {-ft[2], 0.0, ft[0]};
  double fn2[3] = {-ft[1], ft[0], 0.0};
  double lift = 0, drag = 0, latF = 0, pl = 0, pd = 0, fd = 0, latM = 0, longM = 0;
  double pinf = 1. / (P->gamma * P->mach_number * P->mach_number);
  FOR_RANGE(L, -20, ib, ip) {
    double x = L->xyz[ip], y = L->xyz[n + ip], z = L->xyz[2 * (size_t)n + ip];
    double xc = x - P->datum[0], yc = y - P->datum[1], zc = z - P->datum[2];
    if (!(z > P->groundplane + 0.01)) continue;
    double nx = L->nrm[ip], ny = L->nrm[n + ip], nz = L->nrm[2 * (size_t)n + ip];
    double a = L->warea[ip], dp = L->p[ip] - pinf;
    double n1 = nx * fn1[0] + ny * fn1[1] + nz * fn1[2];
    double n2 = nx * fn2[0] + ny * fn2[1] + nz * fn2[2];
    double nt = nx * ft[0] + ny * ft[1] + nz * ft[2];
    lift -= a * dp * n1;
    latF -= a * dp * n2;
    drag -= a * dp * nt;
    latM += xc * a * dp * n2;
    latM -= yc * a * dp * nt;
    longM += xc * a * dp * n1;
    longM -= zc * a * dp * nt;
    pl -= L->p[ip] * (a - pinf) * n1;
    pd -= L->p[ip] * (a - pinf) * nt;
    if (P->reynolds_number > 0.0) {
      double w1 = L->ws[ip], w2 = L->ws[L->nb + ip], w3 = L->ws[2 * (size_t)L->nb + ip];
      double s1 = w1 * fn1[0] + w2 * fn1[1] + w3 * fn1[2];
      double s2 = w1 * fn2[0] + w2 * fn2[1] + w3 * fn2[2];
      double st = w1 * ft[0] + w2 * ft[1] + w3 * ft[2];
      lift += a * s1;
      latF += a * s2;
      drag += a * st;
      fd += a * st;
      latM -= xc * a * s2;
      latM += yc * a * st;
      longM -= xc * a * s1;
      longM += zc * a * st;
    }
  }
  double f = 2.0 / P->reference_area;
  forces[0] = f * lift;
  forces[1] = f * drag;
  forces[2] = f * latF;
  forces[3] = f * fd;
  forces[4] = latM;
  forces[5] = longM;
  forces[6] = f * pl;
  forces[7] = f * pd;
  forces[8] = 0.0;
  for (int j = 0; j < nres; ++j)
    if (res[j] != res[j]) return fail(FLITE_E_NAN, "NaN in residual");
  return 0;
}

/* ------------------------------------------------------------------------
 * field access and single operations
 * ------------------------------------------------------------------------ */

static double *field_ptr(orl *L, int field, int *ncomp) {
  switch (field) {
    case FLITE_FIELD_U: *ncomp = 6; return L->u;
    case FLITE_FIELD_UPREV: *ncomp = 6; return L->uprev;
    case FLITE_FIELD_RHS: *ncomp = 6; return L->rhs;
    case FLITE_FIELD_DISSIPATION: *ncomp = 5; return L->diss;
    case FLITE_FIELD_PRESSURE: *ncomp = 1; return L->p;
    case FLITE_FIELD_DT: *ncomp = 1; return L->dt;
    case FLITE_FIELD_DT_TURB: *ncomp = 1; return L->dtt;
    case FLITE_FIELD_LAMINAR_VISCOSITY: *ncomp = 1; return L->mu;
    case FLITE_FIELD_VORTICITY: *ncomp = 1; return L->vort;
    case FLITE_FIELD_DIVERGENCE: *ncomp = 1; return L->divg;
    case FLITE_FIELD_SOURCE: *ncomp = 6; return L->src;
    case FLITE_FIELD_TURB_DIFFUSION: *ncomp = 1; return L->tdiff;
    case FLITE_FIELD_UOLD: *ncomp = 6; return L->uold;
    case FLITE_FIELD_UOLD2: *ncomp = 6; return L->uold2;
    default: return NULL;
  }
}

int or_field_get(or_ctx *C, int level, int field, double *out, int ld) {
  settle(C);
  if (C && C->P.flow_type == 0 && (field == FLITE_FIELD_UOLD || field == FLITE_FIELD_UOLD2))
    return fail(FLITE_E_STATE, "no solutions of earlier physical times in a steady run");
  if (!C || level < 1 || level > C->nl) return fail(FLITE_E_ARG, "bad level");
  orl *L = &C->L[level - 1];
  if (ld < L->n) return fail(FLITE_E_ARG, "ld < n");
  if (field == FLITE_FIELD_GRADIENT) {
    for (int c = 0; c < 25; ++c)
      for (int i = 0; i < L->n; ++i) out[(size_t)c * ld + i] = H(C, c, i);
    return 0;
  }
  if (field == FLITE_FIELD_WALL_STRESS) {
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < L->n; ++i) out[(size_t)c * ld + i] = i < L->nb ? L->ws[(size_t)c * L->nb + i] : 0.0;
    return 0;
  }
  int nc;
  double *p = field_ptr(L, field, &nc);
  if (!p) return fail(FLITE_E_ARG, "bad field");
  for (int c = 0; c < nc; ++c) memcpy(out + (size_t)c * ld, p + (size_t)c * L->n, sizeof(double) * L->n);
  return 0;
}

int or_field_set(or_ctx *C, int level, int field, const double *in, int ld) {
  settle(C);
  if (C && C->P.flow_type == 0 && (field == FLITE_FIELD_UOLD || field == FLITE_FIELD_UOLD2))
    return fail(FLITE_E_STATE, "no solutions of earlier physical times in a steady run");
  if (!C || level < 1 || level > C->nl) return fail(FLITE_E_ARG, "bad level");
  orl *L = &C->L[level - 1];
  if (ld < L->n) return fail(FLITE_E_ARG, "ld < n");
  int nc;
  double *p = field_ptr(L, field, &nc);
  if (!p) return fail(FLITE_E_ARG, "bad field");
  for (int c = 0; c < nc; ++c) memcpy(p + (size_t)c * L->n, in + (size_t)c * ld, sizeof(double) * L->n);
  return 0;
}

/* per node and component, the sum of the magnitudes of the terms one
 * operation (FLITE_OP_RHS, VISCOUS or SA_DIFFUSION) adds up at the current
 * state: the parity tests' per-node normaliser sum_edges |f_edge| */
int or_magnitudes(or_ctx *C, int level, int op, double *out, int ld) {
  settle(C);
  if (!C || level < 1 || level > C->nl) return fail(FLITE_E_ARG, "bad level");
  if (op != FLITE_OP_RHS && op != FLITE_OP_VISCOUS && op != FLITE_OP_SA_DIFFUSION)
    return fail(FLITE_E_ARG, "magnitudes of RHS, VISCOUS or SA_DIFFUSION only");
  orl *L = &C->L[level - 1];
  if (ld < L->n) return fail(FLITE_E_ARG, "ld < n");
  C->mag = zalloc(6 * (size_t)L->n);
  int rc = or_apply(C, level, op);
  for (int c = 0; c < 6 && !rc; ++c) memcpy(out + (size_t)c * ld, C->mag + (size_t)c * L->n, sizeof(double) * L->n);
  free(C->mag);
  C->mag = NULL;
  return rc;
}

int or_set_threads(or_ctx *C, int threads) {
  if (!C || threads < 1) return fail(FLITE_E_ARG, "threads >= 1");
  C->threads = threads;
  return 0;
}

int or_apply(or_ctx *C, int level, int op) {
  settle(C);
  if (!C || level < 1 || level > C->nl) return fail(FLITE_E_ARG, "bad level");
  orl *L = &C->L[level - 1];
  switch (op) {
    case FLITE_OP_PRESSURE: pressure(C, L, L->u); return 0;
    case FLITE_OP_PRESSURE_ABS:
      pressure(C, L, L->u);
      for (int i = 0; i < L->n; ++i) L->p[i] = fabs(L->p[i]);
      return 0;
    case FLITE_OP_TIMESTEPS: timesteps(C, L); return 0;
    case FLITE_OP_DISSIPATION: return dissipation2(C, L);
    case FLITE_OP_VISCOUS: viscous(C, L, L->diss); return 0;
    case FLITE_OP_RHS: {
      memset(L->rhs, 0, sizeof(double) * 6 * (size_t)L->n);
      int rc = make_rhs(C, L);
      if (rc) return rc;
      if (C->P.turbulence_model == 1) {
        sa_rhs(C, L);
        sa_source(C, L);
      }
      return 0;
    }
    case FLITE_OP_SA_TIMESTEPS: sa_timesteps(C, L); return 0;
    case FLITE_OP_SA_DIFFUSION: sa_diffusion(C, L); return 0;
    default: return fail(FLITE_E_ARG, "bad op");
  }
}

/*
 * flite3d_amd.h -- C-ABI of the MI355X flow-solver hot path.
 *
 * Drop-in for the Fortran module-procedure seam between FLITE3D's
 * MultigridSolver driver (SolverSource/mgnsg3d_mg_par.F90) and its
 * GridSolver module (SolverSource/mgnsg3d_def_par.F90).  A Fortran host binds
 * these entry points through ISO_C_BINDING (flite3d_amd/fortran/flite_gpu.f90)
 * and keeps run.inp, <case>.inp, <case>.sol_1 and the .res/.rsd outputs.
 *
 * Conventions
 *   - plain pointers and sizes, no library types in any signature;
 *   - node-based arrays use the reference layout: column-major SoA with
 *     leading dimension N (u(N,6) is u[c*N + i]);
 *   - index arrays are 0- or 1-based as flite_level_view.index_base says;
 *   - every call is synchronous with respect to the calling host thread and
 *     returns 0 on success or a negative FLITE_E_* code; flite_last_error()
 *     returns the message of the last failure on the calling thread;
 *   - one HIP stream per context; contexts are not shared between threads.
 */
#ifndef FLITE3D_AMD_H
#define FLITE3D_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FLITE_OK = 0,
  FLITE_E_ARG = -1,      /* bad argument / shape */
  FLITE_E_HIP = -2,      /* HIP runtime error */
  FLITE_E_NAN = -3,      /* NaN in the monitor residual (reference STOPs, mgnsg3d_mg_par.F90:1464-1470) */
  FLITE_E_UNSUPPORTED = -4, /* option outside the implemented path */
  FLITE_E_STATE = -5     /* call order (e.g. level not uploaded) */
};

/* Scalars of InputVariablesData (mgnsg3d_def_par.F90:64-196) used by the
 * hot path, after readInputVariables' post-processing
 * (mgnsg3d_def_par.F90:386-415).  Field names follow the ivd fields. */
typedef struct flite_params {
  double gamma;
  double mach_number;
  double reynolds_number;          /* 0 => Euler */
  double prandtl_number;
  double turbulent_prandtl_number;
  double alpha, beta;              /* degrees */
  double inflow_field[2][6];       /* ivd%inflowField(j,k) at [k-1][j-1] */
  double inflow_temperature;
  int walls_are_isentropic;
  int turbulence_model;            /* 0 none, 1 Spalart-Allmaras */
  int flux_type;                   /* 1 central, 2 LLF, 3 Roe, 4 HLLC (DEF:1664-2100) */
  int hllc_flux_order;             /* 1 or 2 */
  int automatic_clipping;          /* 1 = LED local clipping */
  int gradient_method;             /* 1 = Green-Gauss (Galerkin) */
  double hllc_clipping;
  int boundary_term;               /* 1 = FE (3:1), 2 = FV */
  double cfl_number;
  double turbulent_cfl_number;
  double rk_coefficients[5];       /* stage factors alpha_k: 0.6, 0.6, 1.0 (mgnsg3d_def_par.F90:14240-14244) */
  int rk_stages;                   /* 3 as the reference; up to 5 (e.g. 1/4, 1/6, 3/8, 1/2, 1) */
  double max_increment_factor;
  double minimum_density;
  double max_turbulence_value;
  double turb_k;
  double second_order_dissipation_factor;
  double fourth_order_dissipation_factor;
  int multigrid_scheme;            /* 1 V-cycle, 3 growing V */
  int number_of_relaxation_steps;
  int number_of_turbulence_grids;
  int number_of_turbulence_steps;
  double prolongation_relaxation;
  double turb_prolongation_relaxation;
  double multigrid_bc_relaxation;
  double reference_area;
  double groundplane;
  double datum[3];
  int coarse_fp32;                 /* 1: levels >= 2 smooth and transfer in FP32 */
  /* artificial dissipation of the non-HLLC fluxes (makeArtificialDissipation2,
   * mgnsg3d_def_par.F90:11484-11993; defaults DEF:263-267, 316) */
  double coarse_grid_dissipation_factor;  /* fourth-order factor on levels >= 2 */
  int dissipation_scheme;                 /* 2 = JST (scheme 2 only) */
  int coarse_grid_dissipation_scheme;     /* 2 = JST (scheme 2 only) */
  int use_matrix_dissipation;             /* 1: Roe-matrix scaling instead of the scalar d0 */
  double hartens_correction_factor;       /* entropy fix of the matrix scaling */
  /* dual time stepping (addTimeDifferences, mgnsg3d_def_par.F90:1184-1241;
   * solveSystem, 1427-1436) */
  int flow_type;                   /* 0 steady, 1 fixed mesh, 2 moving mesh (flite_level_geometry) */
  int time_scheme;                 /* 2: BDF2 on level 1, otherwise BDF1 */
  int add_time_step;               /* 0 drops the time differences (ivd%addTimeStep) */
  double physical_time_step;
} flite_params;

/* Non-owning view of one level as read from <case>.sol_1
 * (readGridComputationData, mgnsg3d_def_par.F90:13812-14708).  The library
 * copies, renumbers and keeps its own device copies. */
typedef struct flite_level_view {
  int index_base;                  /* 0 (C) or 1 (Fortran) */
  int n_nodes;
  int n_sides;
  int n_bsides;
  int n_bnodes;
  const int *side_index;           /* sideIndexArray(E,2), column-major */
  const double *side_weights;      /* sideWeightsArray(E,3) */
  const int *bside_index;          /* brp%sideIndexArray(Eb,3): i1,i2,indicator */
  const double *bside_weights;     /* brp%sideWeightsArray(Eb,3) */
  const double *node_volume;       /* nodeVolume(N) */
  const double *coordinates;       /* coordinates(N,3); NULL on coarse levels (treated as 0) */
  const double *wall_distance;     /* wallDistance(N) */
  const int *node_indicator_register; /* nodeIndicatorRegister(-20:0), 21 ints, 1-based ranges */
  const int *node_indicator_array; /* nodeIndicatorArray(Nb) */
  /* coarse levels only (NULL on level 1) */
  int n_fine;                      /* nodes of the next finer level */
  const int *prolongation;         /* pod%prolongationArray(n_fine): fine -> coarse */
  const double *restriction;       /* rod%restrictionArray(n_fine): V_f / V_c */
  /* moving meshes (flow_type 2), as readGridComputationData holds them after
   * its read; NULL on a fixed mesh (zero GCL weights and movement, previous
   * volumes equal to nodeVolume) */
  const double *side_gcl;          /* GCLWeightsArray(E): 1.5 g - 0.5 g_prev on level 1
                                      with timeScheme 2 (DEF:13855-13861), g on coarse levels */
  const double *bside_gcl;         /* brp%GCLWeightsArray(Eb,2), the same rule (DEF:816-819) */
  const double *node_volume_p;     /* nodeVolumeP(N), level 1 (DEF:13891) */
  const double *node_volume_p2;    /* nodeVolumeP2(N), level 1 */
  const double *coordinate_movement; /* coordinateMovement(N,3), level 1 (DEF:13896-13900) */
} flite_level_view;

typedef struct flite_ctx flite_ctx;

/* ---- lifecycle -------------------------------------------------------- */
/* Replaces the allocation half of readComputationData (mgnsg3d_mg_par.F90:1198-1268). */
/* n_batch >= 1 flow conditions share the mesh: p points to n_batch structs
 * that may differ only in mach_number, alpha, beta and inflow_field[0]
 * (config of many freestream conditions on one mesh; the reference runs
 * one condition per process).  Every stepping call advances all conditions
 * in the same launches; field, state and monitor calls act on the condition
 * chosen by flite_batch_select (or the batch argument). */
int flite_ctx_create(const flite_params *p, int n_levels, int n_batch, flite_ctx **out);
/* Replaces readGridComputationData's device side (mgnsg3d_def_par.F90:13812-14708). */
int flite_level_upload(flite_ctx *ctx, int level, const flite_level_view *v);
/* The next physical step's mesh of a moving-mesh run (flow_type 2): the
 * readComputationData(COMPINFILE, 1) of multiGridCycles (mgnsg3d_mg_par.F90:
 * 525-548), which replaces the geometry and keeps every solution array.
 * Same node, side and boundary-side counts as the level's last upload; u,
 * uold, uold2 and the laminar viscosity of every level and condition carry
 * over.  Call it for each level whose geometry changes, finest first.  The
 * next operation then runs the step's preamble (mgnsg3d_mg_par.F90:586-601)
 * on the new geometry: laminar viscosity reset, u and uold restricted to
 * every coarse level. */
int flite_level_geometry(flite_ctx *ctx, int level, const flite_level_view *v);
void flite_ctx_destroy(flite_ctx *ctx);
const char *flite_last_error(void);

/* ---- state ------------------------------------------------------------ */
/* Solution u(N,ncomp) in host numbering, leading dimension ld >= N. */
int flite_state_set(flite_ctx *ctx, int level, const double *u, int ld, int ncomp, int batch);
int flite_state_get(flite_ctx *ctx, int level, double *u, int ld, int ncomp, int batch);
/* Freestream start (setUpInitialField2, mgnsg3d_def_par.F90:3713-3996) on
 * level 1 followed by the cycle preamble of multiGridCycles
 * (mgnsg3d_mg_par.F90:586-602): laminar viscosity and restricted u on
 * every level. */
int flite_init_freestream(flite_ctx *ctx);
/* Same preamble when u on level 1 came from flite_state_set (restart). */
int flite_init_from_state(flite_ctx *ctx);

/* ---- stepping (mgnsg3d_mg_par.F90:939-1186 callers) --------------------- */
/* End of a physical time step of an unsteady run (flow_type 1 or 2):
 * uold2 <- uold <- u on level 1, as multiGridCycles does after each
 * physical step (mgnsg3d_mg_par.F90:889-890), then the next step's preamble,
 * which the reference runs inside its physical-step loop
 * (mgnsg3d_mg_par.F90:586-601): laminar viscosity reset, u and uold
 * restricted to every coarse level. */
int flite_advance_physical_time(flite_ctx *ctx);
/* doTimeIterations (mgnsg3d_def_par.F90:966-1000) */
int flite_do_time_iterations(flite_ctx *ctx, int level, int relaxation_steps,
                             int calculate_dissipation, int calculate_turbulence,
                             int iteration_number);
/* restrictGrid (mgnsg3d_def_par.F90:12141-12430), fine_level -> fine_level+1 */
int flite_restrict(flite_ctx *ctx, int fine_level, int calculate_turbulence);
/* prolongateGrid (mgnsg3d_def_par.F90:12432-12570), coarse_level -> coarse_level-1 */
int flite_prolongate(flite_ctx *ctx, int coarse_level);
/* mapFineToCoarse of u (mgnsg3d_def_par.F90:12572-12690), as oneCycle calls it */
int flite_map_u_fine_to_coarse(flite_ctx *ctx, int fine_level);
/* oneCycle (mgnsg3d_mg_par.F90:907-1196), schemes 0, 1 and 3 */
int flite_cycle(flite_ctx *ctx, int cycle_number);

/* ---- monitor ------------------------------------------------------------ */
/* writeResiduals' reductions (mgnsg3d_mg_par.F90:1271-1445) and
 * calculateLiftAndDrag (mgnsg3d_def_par.F90:14710-14870) on level 1:
 *   res[7]     sqrt(sum (u_j - uprev_j)^2), j < nres (5, or 6 with SA)
 *   maxres[7]  max |u_j - uprev_j|, maxidx[7] its node (host numbering, index_base)
 *   forces[9]  lift, drag, latF, frictionDrag, latM, longM, pressureLift, pressureDrag, 0 */
int flite_monitor(flite_ctx *ctx, double res[7], double maxres[7], int maxidx[7], double forces[9]);
/* flite_monitor without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as flite_monitor, with the block records in a
 * file-scope device array of the library (119 KB per GPU, held until the
 * process ends) instead of the context's buffer; allocates nothing, host or
 * device.  The context must be built (any earlier stepping, field or monitor
 * call builds it; FLITE_E_STATE otherwise).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two host threads, two
 * contexts' streams) race. */
int flite_monitor_noalloc(flite_ctx *ctx, double res[7], double maxres[7], int maxidx[7], double forces[9]);

/* ---- parity / inspection ------------------------------------------------ */
/* Named node fields of a level, host numbering, column-major (N, ncomp):
 *   FLITE_FIELD_U (6), UPREV (6), RHS (6), DISSIPATION (5), PRESSURE, DT,
 *   DT_TURB, LAMINAR_VISCOSITY, VORTICITY, DIVERGENCE, SOURCE (6),
 *   TURB_DIFFUSION, GRADIENT (25: nodeHelpArray 1-15 gradients, 16-25 limits),
 *   WALL_STRESS (3, boundary nodes only, zeros elsewhere),
 *   UOLD (6), UOLD2 (6): solutions at the last two physical times (grp%uold,
 *   grp%uold2; unsteady runs only) */
enum {
  FLITE_FIELD_U = 0, FLITE_FIELD_UPREV, FLITE_FIELD_RHS, FLITE_FIELD_DISSIPATION,
  FLITE_FIELD_PRESSURE, FLITE_FIELD_DT, FLITE_FIELD_DT_TURB,
  FLITE_FIELD_LAMINAR_VISCOSITY, FLITE_FIELD_VORTICITY, FLITE_FIELD_DIVERGENCE,
  FLITE_FIELD_SOURCE, FLITE_FIELD_TURB_DIFFUSION, FLITE_FIELD_GRADIENT,
  FLITE_FIELD_WALL_STRESS, FLITE_FIELD_UOLD, FLITE_FIELD_UOLD2, FLITE_FIELD_COUNT
};
int flite_field_get(flite_ctx *ctx, int level, int field, double *out, int ld);
int flite_field_set(flite_ctx *ctx, int level, int field, const double *in, int ld);
int flite_field_ncomp(int field);

/* Single operations of the smoother, for parity tests against the CPU
 * restatement; each reads and writes the level state like its reference
 * routine. */
enum {
  FLITE_OP_PRESSURE = 0,      /* makePressureField (13612) */
  FLITE_OP_PRESSURE_ABS,      /* makePressureField; p = |p| (1024-1025) */
  FLITE_OP_TIMESTEPS,         /* makeTimeSteps (13335) */
  FLITE_OP_DISSIPATION,       /* makeArtificialDissipation2 early-return semantics (11484) */
  FLITE_OP_VISCOUS,           /* makeViscosityTerm into dissipation (5198) */
  FLITE_OP_RHS,               /* rhs = 0; makeRHS (1525); with SA also
                                 makeSARHS + makeSASourceTerm (6809, 9285) */
  FLITE_OP_SA_TIMESTEPS,      /* makeTurbulentTimeSteps (12878) */
  FLITE_OP_SA_DIFFUSION,      /* makeSADiffusion (7436) */
  FLITE_OP_COUNT
};
int flite_apply(flite_ctx *ctx, int level, int op);

/* ---- linearisation (discrete tangent and adjoint) ---------------------- */
/* The level-1 convective residual R(U) of makeRHS (mgnsg3d_def_par.F90:
 * 1525-2512: interior HLLC sides and boundary sides, as FLITE_OP_RHS
 * computes it) differentiated exactly at the current state, for first- and
 * second-order HLLC (MUSCL reconstruction with the Green-Gauss gradients and
 * LED clipping), inviscid or laminar: with ReynoldsNumber > 0 the residual
 * includes the viscous term of makeViscosityTerm (mgnsg3d_def_par.F90:
 * 5198-5893, the dissipation the smoother adds).  With Spalart-Allmaras
 * (turbulence_model 1) the tangent and the adjoint also differentiate the
 * eddy viscosity in the stress (DEF:5556-5566), the nu~ convection
 * (makeSARHS, DEF:6809-6950), the SA diffusion (DEF:7436-7808) and the
 * source (DEF:9285-9370): du, lam, out and dJdU are then (N, 6), nu~ last.  The
 * second-order residual is piecewise smooth: at a limiter switch these give
 * the derivative of the branch the state selects.  Vectors are otherwise
 * (N, 5) column-major in host numbering with leading dimension ld.  The reference has no differentiated code; these entry
 * points serve the sensitivity calculations a maintainer adds on top of it.
 * Both recompute the level's pressure, gather records and (second order)
 * gradients from the current state.
 *   flite_tangent_apply:  out = (dR/dU) du
 *   flite_adjoint_apply:  out = (dR/dU)^T lam  (a reversed side loop that
 *                         re-evaluates each side's inputs on dual numbers;
 *                         no Jacobian is stored)
 *   flite_drag_state_gradient: the drag of calculateLiftAndDrag
 *                         (mgnsg3d_def_par.F90:14757-14803: the pressure
 *                         term and, with ReynoldsNumber > 0, the friction
 *                         term A_i ws_i.t of the wall stress, recomputed at
 *                         the current state) and its partial derivative
 *                         dJ/dU. */
int flite_tangent_apply(flite_ctx *ctx, const double *du, int ld_du, double *out, int ld_out);
int flite_adjoint_apply(flite_ctx *ctx, const double *lam, int ld_lam, double *out, int ld_out);
int flite_drag_state_gradient(flite_ctx *ctx, double *drag, double *dJdU, int ld);
/* The tangent and adjoint with the vectors in device memory on the
 * context's GPU (no counterpart in the reference; for solvers that keep
 * their vectors in HBM and iterate the applies without host copies): du,
 * lam and out are (nc, N) contiguous, host numbering, nc = 5 (6 with
 * Spalart-Allmaras).  The call waits for all earlier work on the device
 * (a producer on another stream has then finished), and returns when out
 * is written.  Host pointers return FLITE_E_ARG. */
int flite_tangent_apply_device(flite_ctx *ctx, const double *du, double *out);
int flite_adjoint_apply_device(flite_ctx *ctx, const double *lam, double *out);

/* Derivatives with respect to the level-1 mesh metrics, for gradients with
 * respect to the node coordinates (no counterpart in the reference; for
 * shape-sensitivity callers).  Same residual and restrictions as
 * flite_adjoint_apply, first-order HLLC and inviscid flow only.  Outputs use the uploaded side order and layout,
 * column-major (E,3) like flite_level_view.side_weights and (Eb,3) like
 * bside_weights, with leading dimensions ld_w >= n_sides, ld_wb >= n_bsides:
 *   flite_metric_adjoint:  gw  = d(lam^T R)/d side_weights,
 *                          gwb = d(lam^T R)/d bside_weights, at the current
 *                          state (dual numbers seeded on the weights);
 *   flite_drag_metric_gradient: the drag of an inviscid run (its pressure term)
 *                          and its partial derivative with respect to the
 *                          boundary-side weights, through the wall-node areas
 *                          and normals the library forms from them.
 * With lam solving (dR/dU)^T lam = dJ/dU, the coordinate gradient of J is
 * the chain of (gJwb - gwb, -gw) through the metrics' dependence on the
 * coordinates (fp_metric_vjp in flite3d_prepro.h). */
int flite_metric_adjoint(flite_ctx *ctx, const double *lam, int ld_lam, double *gw, int ld_w, double *gwb, int ld_wb);
int flite_drag_metric_gradient(flite_ctx *ctx, double *drag, double *gwb, int ld_wb);
/* The same for first- and second-order HLLC, inviscid, laminar or
 * Spalart-Allmaras (no counterpart in the reference; for shape-sensitivity
 * callers): d(lam^T R) with respect to every mesh quantity the level-1
 * residual of flite_adjoint_apply reads,
 *   gw, gwb   side and boundary-side weights (layout as flite_metric_adjoint),
 *   gvol[N]   dual volumes (host numbering),
 *   gxyz      coordinates directly (N,3) column-major, ld_x >= N,
 *   gwd[N]    wall distances (host numbering; Spalart-Allmaras only, where
 *             the source reads them; may be NULL otherwise).
 * Second order reads the volumes and the weights through the Green-Gauss
 * gradients and the coordinates through the reconstruction's side vector
 * (x_b - x_a) and clip length; first order gives gxyz = 0.  The viscous
 * term reads the weights and volumes through the stress's Green-Gauss
 * gradients, the weights in its side and boundary fluxes and the wall
 * normals in the adiabatic-wall projection (their shares reach gwb through
 * the normals' sums of boundary-side weights); the SA terms read the same
 * through the nu~ gradients, the diffusion's side terms and the source.
 * lam is (N, 6) with Spalart-Allmaras.
 *   flite_drag_mesh_gradient: the drag (as flite_drag_state_gradient, with
 *             friction when ReynoldsNumber > 0) and its partial derivatives
 *             with respect to the side weights, boundary-side weights and
 *             volumes (the friction's wall stress reads all three; the wall
 *             areas and normals read the boundary-side weights).
 * The coordinate gradient of J is then fp_mesh_vjp's chain of
 * (gJw - gw, gJwb - gwb, gJvol - gvol, -gxyz) (flite3d_prepro.h), plus
 * -gwd through the wall distance |x_i - x_w(i)| to the nearest wall node. */
int flite_mesh_adjoint(flite_ctx *ctx, const double *lam, int ld_lam, double *gw, int ld_w, double *gwb, int ld_wb,
                       double *gvol, double *gxyz, int ld_x, double *gwd);
int flite_drag_mesh_gradient(flite_ctx *ctx, double *drag, double *gw, int ld_w, double *gwb, int ld_wb,
                             double *gvol);

/* Precision of the coarse levels from the next call on: 1 = FP32, 0 = FP64
 * (flite_params.coarse_fp32 sets the initial one).  No counterpart in the
 * reference.  FP32 coarse corrections are cheaper, but the FAS forcing
 * on a coarse level is the difference of O(1) residuals, rounded to FP32, so
 * the fine residual stalls near 1e-5 of its start; a driver that needs a
 * deeper drop runs FP32 until the stall and then switches to FP64, which
 * converges to the FP64 solution.  The first switch builds the other
 * precision's coarse levels (as at upload); later switches swap them.  The
 * coarse states are rebuilt from the fine state at the next restriction. */
int flite_set_coarse_precision(flite_ctx *ctx, int fp32);

/* Autotuning of the launch shape (no counterpart in the reference): per
 * level, the lanes per node (1, 4 or 16) of the gather and stage kernels
 * and the side-block shape of the flux kernel (256 sides touching at most
 * 128 nodes, or 512 touching at most 224) that run fastest on the
 * context's current state, each candidate timed with HIP events over
 * `repeats` runs of the kernels it changes.  The choice holds for the
 * context's life; the lanes are written to lanes_out[0 .. n_levels-1].
 * Without autotuning a level gets 1 lane from 2^20 nodes, 4 from 2^16,
 * else 16 (FLITE_GATHER_LANES forces one value), and 256-side blocks
 * (FLITE_SIDE_BLOCK=256|512 forces a shape).  Lanes and shapes change the
 * order of the gather and node sums (results agree to rounding), and
 * timing noise can change the choice from run to run.  flite_level_lanes
 * and flite_level_side_block read a level's current setting (sides per
 * block and its node cap). */
int flite_autotune(flite_ctx *ctx, int repeats, int *lanes_out);
int flite_level_lanes(flite_ctx *ctx, int level, int *lanes);
int flite_level_side_block(flite_ctx *ctx, int level, int *sides, int *nodes);

/* Condition that flite_field_get/set and flite_monitor act on (default 0);
 * no reference counterpart (the reference has one condition per run). */
int flite_batch_select(flite_ctx *ctx, int batch);

/* Wait for all work queued by the context. */
int flite_sync(flite_ctx *ctx);

/* Kernel timing with HIP events recorded on the context's stream around
 * every launch of the named kernel class (e.g. "flux_l1", "stage_l1",
 * "grad_l1"; "_lN" = level N).  flite_timing_enable(ctx, name) starts
 * recording (name NULL stops); flite_timing_get returns the summed
 * duration (ms) and the number of launches since enabling. */
int flite_timing_enable(flite_ctx *ctx, const char *name);
int flite_timing_get(flite_ctx *ctx, double *total_ms, int *launches);
/* Per-cycle counts of side-loop evaluations: writes, per level, the number
 * of residual (flux) evaluations one flite_cycle performs. */
int flite_cycle_evaluations(flite_ctx *ctx, int *evals_per_level);
/* Sizes of a level as uploaded: nodes, sides, boundary sides. */
int flite_level_sizes(flite_ctx *ctx, int level, int *n_nodes, int *n_sides, int *n_bsides);

#ifdef __cplusplus
}
#endif
#endif

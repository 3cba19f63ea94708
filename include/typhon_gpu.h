/*
 * typhon_gpu.h -- C-ABI of the MI355X (gfx950) residual path for Typhon.
 *
 * Drop-in boundary for the body of calc_rhs and the RK stage loop
 * (SOURCE/MAIN/calc_rhs.f90:10, SOURCE/MAIN/treelevel_rungekutta.f90:8).  The
 * caller (Typhon's Fortran driver through iso_c_binding, or the Python host
 * mirror in typhon_amd/) owns every host array; the library copies what it
 * needs once and keeps mesh and state resident in HBM.  Plain pointers and
 * sizes only.  Every function returns 0 on success, a TY_ERR_* code otherwise;
 * ty_last_error() gives the message (no exceptions cross the ABI).
 *
 * Threading: one host thread per context, one HIP stream per context.
 */
#ifndef TYPHON_GPU_H
#define TYPHON_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TY_OK 0
#define TY_ERR_ARG 1          /* bad argument or call order */
#define TY_ERR_HIP 2          /* HIP runtime error */
#define TY_ERR_STATE 3        /* non-positive or NaN internal energy (calc_ns_timestep.f90:89-100) */
#define TY_ERR_UNSUPPORTED 4  /* scheme/BC combination not provided by this library */
#define TY_ERR_NOT_CONVERGED 5

#define TY_PREC_FP64 0        /* everything in FP64 */
#define TY_PREC_MIXED 1       /* FP32 stage state, gradients, fluxes, geometry; FP64 U0/accumulator */
#define TY_PREC_MIXED_F32SUM 2 /* TY_PREC_MIXED with the RK stage-increment sum kept in FP32 between
                                  stages (U0 and its update stay FP64): 30 B less per cell and RK4
                                  stage; no counterpart in the reference, which is FP64 throughout */

typedef struct ty_ctx ty_ctx;

/* One boundary condition: umesh%boco(ib) (CFDTOOLS/Mesh/USTMESH.f90:31-52) with
 * defsolver%boco(idef)%typ_boco (SOURCE/MODCOM/VARCOM.f90:75-93) and its
 * resolved uniform parameters (setboco_ns_*.f90). */
typedef struct {
    int32_t type;            /* 10 symmetry, 11 periodic, 20 adiabatic wall, 21 isothermal wall,
                                22 heat-flux wall, 30 subsonic inlet, 31 supersonic inlet,
                                35 subsonic outlet, 36 supersonic outlet */
    int32_t nface;
    const int32_t* iface;    /* umesh%boco(ib)%iface(1:nface), 1-based */
    const int32_t* ipart;    /* periodic only: boco%gridcon%i_param(1:nface), 1-based partner faces */
    double param[8];         /* 31: ptot, ttot, mach, dir_x, dir_y, dir_z; 35: pstat;
                                30: ptot, ttot, unused, dir_x, dir_y, dir_z (setboco_ns_inlet_sub.f90,
                                    uniform values);
                                21: temp_wall, -, -, -, wall velocity x, y, z (setboco_ns_isoth.f90);
                                22: boco_ns%flux (the user's WALL_FLUX negated, def_boco_ns.f90:73-74),
                                    -, -, -, wall velocity x, y, z (setboco_ns_flux.f90, ns_bocoflux.f90);
                                20: -, -, -, -, wall velocity x, y, z (a 22 with zero flux,
                                    def_boco_ns.f90:44-49).  20 and 22 need the laminar NS model.
                                23 CONVECTION_WALL and 24 GEN_CONV_WALL are refused (TY_ERR_UNSUPPORTED):
                                the reference refuses both for NS (MENU_NS.f90:95-113 has no case
                                for them; setboco_ns_hconv.f90:78 stops). */
} ty_boco_desc;

/* The unstructured mesh as Typhon holds it after init_world (SOURCE/MAIN/main.f90:95):
 * st_ustmesh (USTMESH.f90:31-52), st_mesh (MESHBASE.f90:44-59). */
typedef struct {
    int32_t dim;             /* 2 or 3 */
    int32_t ncell_int;       /* umesh%ncell_int */
    int32_t ncell;           /* umesh%ncell (interior + ghost cells) */
    int32_t nface;           /* umesh%nface */
    int32_t nface_int;       /* umesh%nface_int */
    const int32_t* facecell; /* umesh%facecell%fils(1:nface,1:2), column-major, 1-based */
    const double* face;      /* umesh%mesh%iface(1:nface,1,1): st_face {normale(3), centre(3), surface} */
    const double* centre;    /* umesh%mesh%centre(1:ncell,1,1): v3d */
    const double* volume;    /* umesh%mesh%volume(1:ncell,1,1) */
    int32_t nboco;
    const ty_boco_desc* boco;
    int32_t cell_order;      /* device cell order (no reference counterpart): 0 reverse Cuthill-McKee,
                                1 Hilbert curve through the cell centres, 2 the caller's order */
} ty_mesh_desc;

/* st_espece (SOURCE/EQNS/EQNS.f90:60-70) and the fluid model (def_model_ns.f90) */
typedef struct {
    double gamma;
    double r_const;
    double prandtl;
    int32_t typ_visc;        /* 11 Sutherland, 13 constant dynamic, 14 constant kinematic */
    double visc_dyn, visc_kin, tref, tsuth;
    int32_t typ_fluid;       /* 10 Euler, 15 laminar Navier-Stokes (EQNS.f90:27-32) */
} ty_gas;

/* mnu_spat / mnu_time (SOURCE/PARAM/MENU_NUM.f90:155-231) with resolved integer IDs */
typedef struct {
    int32_t sch_hyp;         /* 5 Rusanov, 10 Roe, 21 HLLE, 22 HLLEB, 25 HLLC, 26 HLLCB */
    int32_t sch_dis;         /* 1 none, 5 cell-average, 10 cell-full */
    char method;             /* 'N' none, 'M' MUSCL, 'F' MUSCL-FAST, 'U' MUSCL-UNS */
    char limiter;            /* 'X','M','A','V','S','K','L'; 'W' = Venkatakrishnan cell limiter */
    int32_t gradmeth;        /* 5 none, 10 LSQ, 12 weighted LSQ */
    int32_t postlimiter;     /* 0 none, 20 Barth, 21 super-Barth */
    double venkat_k;         /* Venkatakrishnan K (eps^2 = (K h)^3) */
    int32_t nstage;          /* rk%nstage (1..8) */
    const double* rk_coef;   /* rk%coef(1:nstage,1:nstage), column-major */
    double cfl;              /* deftime%stabnb */
    int32_t precision;       /* TY_PREC_FP64, TY_PREC_MIXED or TY_PREC_MIXED_F32SUM */
} ty_scheme;

/* Matrix-free implicit solve (no reference counterpart: the reference's INVERSION=GMRESFREE
 * path is dead code, SOURCE/MGRID/MATFREE.f90:73-158, gmres_free.f90:147-289).  The linear
 * system of the linearised implicit step is (I - J) x = b with J = dR/dU and
 * R(U) = -dt/V sum_f F.A (SOURCE/MAIN/calc_rhs.f90 convention); J v comes from finite
 * differences of R, the preconditioner is block-Jacobi with the diagonal blocks of the
 * reference's HLL Jacobian assembly (build_implicit_bdlu.f90:44-83). */
typedef struct {
    int32_t nkrylov;         /* GMRES restart length m (1..31) */
    int32_t max_restart;     /* GMRES cycles per refinement pass */
    int32_t max_outer;       /* FP64 refinement passes */
    int32_t inner_precision; /* TY_PREC_MIXED: FP32 Krylov vectors, preconditioner and JVP;
                                TY_PREC_FP64: all FP64 */
    int32_t precond;         /* 0 none, 1 block-Jacobi */
    double inner_tol;        /* relative residual reduction of each GMRES solve */
    double outer_tol;        /* target ||b - (I - J) x|| / ||b|| (FP64 central-difference JVP) */
} ty_krylov;

typedef struct {
    int32_t outer_iters;     /* refinement passes performed */
    int32_t inner_iters;     /* total Arnoldi steps */
    double rel_res;          /* final ||b - (I - J) x|| / ||b|| */
    double rhs_norm;         /* ||b|| */
} ty_krylov_stats;

/* context */
int ty_create(ty_ctx** ctx, int device);
int ty_destroy(ty_ctx* ctx);
const char* ty_last_error(const ty_ctx* ctx);

/* mesh upload once (after init_world); builds the device layout: cells renumbered
 * (mesh->cell_order: reverse Cuthill-McKee or Hilbert, computed on the device),
 * periodic face pairs merged, cell->face slot lists.  Replaces the mesh side of
 * calc_rhs (calc_rhs.f90:48-145). */
int ty_mesh_upload(ty_ctx* ctx, const ty_mesh_desc* mesh);

/* model and scheme: the resolved mnu_solver / mnu_spat / mnu_time menus */
int ty_set_scheme(ty_ctx* ctx, const ty_gas* gas, const ty_scheme* scheme);

/* conservative state of interior cells: etatcons tabscal(1)=rho, tabscal(2)=rhoE,
 * tabvect(1)=momentum v3d (DEFFIELD.f90:32-52); host arrays in the caller's cell order */
int ty_state_set(ty_ctx* ctx, const double* rho, const double* rhoE, const double* mom);
int ty_state_get(ty_ctx* ctx, double* rho, double* rhoE, double* mom);

/* global time step: calc_zonetimestep with stab_cond (ZONE/calc_zonetimestep.f90:57-126,
 * EQNS/calc_ns_timestep.f90:52-102): dt = min(min_c CFL 2V/sum(A)/(|u|+a), dtmax).
 * Returns TY_ERR_STATE on a non-positive or NaN sound speed. */
int ty_timestep(ty_ctx* ctx, double dtmax, double* dt);

/* one residual evaluation, the reference convention -dt/V sum_f F.A on interior cells
 * (calc_rhs.f90:10 + flux_to_res.f90:126-142) for the current state; outputs in the
 * caller's cell order (res_mom is v3d) */
int ty_residual(ty_ctx* ctx, double dt, double* res_rho, double* res_rhoE, double* res_mom);

/* one explicit RK step with a given global dt: treelevel_rungekutta
 * (MAIN/treelevel_rungekutta.f90:50-113) then update_field (ZONE/update_field.f90:38) */
int ty_rk_step(ty_ctx* ctx, double dt);

/* nsteps device-resident steps, each: time step on the device (min with dtmax) then
 * the RK step, no host synchronisation in between (integration_cyclezone.f90:80-151).
 * dt_last, if not NULL, receives the last step's dt (synchronises). */
int ty_advance(ty_ctx* ctx, int32_t nsteps, double dtmax, double* dt_last);

/* wait for queued work; reports a bad state found by device-side time steps */
int ty_sync(ty_ctx* ctx);

/* diagonal blocks I + dt/V sum_f A_f dF_f/dU_c of the implicit operator at the current state
 * (init_implicit_bdlu.f90:37-49 + build_implicit_bdlu.f90:44-83 restricted to the diagonal, HLL
 * face Jacobians calc_jac_hll.f90:46-57 at the face states of the residual).  diag: [ncell_int][5][5]
 * row-major in the caller's cell order, variables (rho, rhoE, rho u, rho v, rho w); in 2D the
 * z-momentum row and column hold the identity. */
int ty_jacobian_diag(ty_ctx* ctx, double dt, double* diag);

/* J v at the current state by central differences of the FP64 residual (v, jv: [5][ncell_int]
 * in the caller's cell order) */
int ty_jvp(ty_ctx* ctx, double dt, const double* v, double* jv);

/* solve (I - J) x = b at the current state (b, x: [5][ncell_int], caller order) by FP64
 * iterative refinement around preconditioned restarted GMRES; TY_ERR_NOT_CONVERGED if the
 * outer target is not reached (x then holds the last iterate) */
int ty_linear_solve(ty_ctx* ctx, double dt, const ty_krylov* kry, const double* b, double* x,
                    ty_krylov_stats* stats);

/* one linearised backward-Euler step: U <- U + x with (I - J) x = R(U) (calc_rhs.f90:133-140
 * tstep_implicit convention) */
int ty_implicit_step(ty_ctx* ctx, double dt, const ty_krylov* kry, ty_krylov_stats* stats);

/* One implicit step of the reference's assembled path (METHOD=IMPLICIT with INVERSION=GMRES):
 * calc_rhs with HLL face Jacobians (calc_jac_hll.f90:46-57, viscous diagonal
 * calc_flux_viscous_cell.f90:149-158, ns_bocojacobian.f90:44-92), the block matrix of
 * build_implicit_bdlu.f90:44-83 over the interior cells and one ghost unknown per boundary face
 * side (ghost rows: identity, right-hand side = the boundary-flux sum flux_to_res leaves there),
 * bdlu_gmres.f90:55-212 (restarted GMRES(nkrylov), modified Gram-Schmidt, no preconditioner,
 * x0 = rhs/diag, restarts while ||r||/||rhs|| >= maxres ||x0||_1 and nit <= 2 max_it; maxres =
 * INV_RES, max_it = IMP_MAX_IT, def_time.f90:204-206), then U += solution on interior cells
 * (tstep_implicit.f90 -> implicit_solve.f90:39-84, update_field.f90:38).  FP64 throughout.
 * info: the GMRES info (iterations - 1, or -1 when it did not converge). */
int ty_implicit_bdlu_step(ty_ctx* ctx, double dt, int32_t nkrylov, double maxres, int32_t max_it, int32_t* info);

/* Tangent-linear residual (no reference counterpart: the reference has no derivative code).
 * Exact forward-mode derivative of the first-order residual R = -dt/V sum_f F.A
 * (calc_rhs.f90:10, flux_to_res.f90:47-142, first-order face states distrib_field.f90) at the
 * current state and a fixed dt:
 *   dres = dR/dU du + dR/dS ds + dR/dV dvol + dR/dq dq
 * du, dres: [5][ncell_int] conservative (rho, rhoE, rho u, rho v, rho w), caller cell order;
 * ds: [nface][3] direction of each face's area vector S = surface * normale (caller faces);
 * dvol: [ncell_int] cell volumes; dq: [nboco][5] ghost primitive state (rho, p, u, v, w) of
 * supersonic inlets (other bocos ignored).  Any input may be NULL (zero direction).
 * Viscous (Navier-Stokes) residuals are covered for du, dvol and dq; their flux also reads
 * cell and face centres, so ds must be NULL for them.
 * TY_ERR_UNSUPPORTED unless the scheme is first order and FP64 (TY_PREC_FP64: a mixed-precision
 * context does not evaluate the FP64 residual this differentiates). */
int ty_residual_tangent(ty_ctx* ctx, double dt, const double* du, const double* ds, const double* dvol,
                        const double* dq, double* dres);
/* Adjoint residual (no reference counterpart): the transpose of ty_residual_tangent,
 * gu = (dR/dU)^T lam, gs = (dR/dS)^T lam, gvol = (dR/dV)^T lam, gq = (dR/dq)^T lam, same
 * layouts; gs, gvol and gq may be NULL.  Far faces of merged periodic pairs get zero in gs
 * (the device uses the geometry of the pair's first face).  Viscous residuals: HLLC/HLLCB
 * fluxes, gs must be NULL.  FP64 schemes only, as ty_residual_tangent. */
int ty_residual_adjoint(ty_ctx* ctx, double dt, const double* lam, double* gu, double* gs, double* gvol,
                        double* gq);

typedef struct ty_unsteady_stats {
    int32_t forward_steps;  /* RK steps advanced for checkpoints and recomputation */
    int32_t adjoint_steps;  /* RK steps taken in reverse (= nsteps) */
    int32_t max_snapshots;  /* checkpoint states held at once (besides the initial state) */
    int32_t snapshots_allocated; /* checkpoint buffers allocated: min(snaps, nsteps - 1) */
    double ms;              /* wall time of the call */
} ty_unsteady_stats;

/* Unsteady adjoint (no reference counterpart): the gradient of an objective J(U_N) of the state
 * after nsteps RK steps of the current scheme (tableau of ty_set_scheme, treelevel_rungekutta.f90:
 * 50-113) from the current state U_0 with the time steps dts[nsteps], given lamN = dJ/dU_N
 * ([5][ncell_int], caller order).  Returns lam0 = dJ/dU_0 (same layout) and, when gq is not
 * NULL, dJ/dq ([nboco][5], supersonic-inlet ghost states (rho, p, u, v, w), summed over all
 * stages and steps).  Checkpoints live in device memory: at most `snaps` states besides U_0,
 * placed by binomial checkpointing (Griewank's treeverse), so the forward work grows as
 * nsteps * r with C(snaps + r, snaps) >= nsteps.  The gradient is that of the FP64 discrete map:
 * the forward recomputation and every reversed step run in FP64, also in a mixed-precision
 * context (which then holds two extra FP64 stage-state buffers for the call).  The state is U_0
 * again on return.  Same scheme and boundary limits as ty_residual_adjoint, any precision. */
int ty_unsteady_adjoint(ty_ctx* ctx, int32_t nsteps, const double* dts, int32_t snaps, const double* lamN,
                        double* lam0, double* gq, ty_unsteady_stats* stats);
/* Steady mode (time_steady).  Reference values of the conserved variables: calc_refcons
 * (SOURCE/ZONE/calc_refcons.f90:20-117) with define_refcons (SOURCE/PARAM/MENU_SOLVER.f90:262-276),
 * out[3] = volume averages of rho and rhoE over the interior cells of the current state, and the
 * momentum reference sqrt(out[0] out[1]) (defsolver%refsca(1:2), defsolver%refvec(1)). */
int ty_refcons(ty_ctx* ctx, double* out);

/* per-cell time steps of loc_stab_cond at the scheme's CFL number (calc_zonetimestep.f90:64-66 with
 * calc_ns_timestep.f90:52-102, not reset to the minimum): dt_cells[ncell_int] in the caller's cell
 * order, dt_min (may be NULL) their minimum.  TY_ERR_STATE on a bad state. */
int ty_local_timestep(ty_ctx* ctx, double* dt_cells, double* dt_min);

/* One steady cycle of integration_cyclezone (SOURCE/MAIN/integration_cyclezone.f90:80-167). */
typedef struct {
    int32_t local_dt;        /* 1: loc_stab_cond (each cell's own time step), 0: stab_cond (global minimum) */
    int32_t maxit;           /* iterations of the cycle at most (zinfo%maxit, integration_cycle.f90:92) */
    int32_t check_every;     /* iterations between the host's looks at the device's stop flag (>= 1) */
    double cfl;              /* deftime%stabnb */
    double cfl_max;          /* deftime%stabnb_max (CFL_MAX, def_time.f90:73) */
    double maxres;           /* zone convergence cur_res / residu_ref <= maxres (check_end_cycle.f90:33) */
    double residu_ref;       /* the zone's reference residual at the cycle start (integration_cycle.f90:94) */
    double wlog;             /* log10(wcur_res / wres_ref) of the world (calc_zonetimestep.f90:70-72; 0 for one
                                zone in its first cycle) */
    double refsca[2];        /* defsolver%refsca (rho, rhoE), from ty_refcons at the initial state */
    double refvec;           /* defsolver%refvec */
} ty_steady;

typedef struct {
    int32_t iterations;      /* iterations performed */
    int32_t converged;       /* cur_res / residu_ref <= maxres at the end */
    double cur_res;          /* update_field's residual of the last iteration */
    double residu_ref;       /* the zone's reference residual at the end */
    double cfl;              /* CFL number the next iteration would use */
} ty_steady_stats;

/* Per iteration, all on the device: the time step with the steady CFL ramp
 * cfl = min(stabnb max(1, 1 - log10(cur_res / residu_ref) - wlog), stabnb_max) (calc_zonetimestep.f90:
 * 64-73), the scheme's RK step with the local or global time step (treelevel_rungekutta.f90:50-113),
 * update_field's residual (update_field.f90:42-60: sum over interior cells of |residu|/refsca for rho
 * and rhoE and |residu momentum|/refvec, residu = the step's increment), check_end_cycle
 * (check_end_cycle.f90:32-33: residu_ref = max(residu_ref, cur_res), stop at cur_res/residu_ref <= maxres
 * or maxit iterations).  res_hist[maxit] (may be NULL) receives cur_res of every iteration (the monitor
 * of write_monitors_iteration.f90:32-35).  The host synchronises only every check_every iterations;
 * launches after the stop do nothing, so the iteration count is the reference's. */
int ty_steady_cycle(ty_ctx* ctx, const ty_steady* steady, double* res_hist, ty_steady_stats* stats);

/* Volume probes (BLOCK:PROBE, def_probe.f90:34-80; the reference's codes, DEFPROBE.f90:26-28). */
#define TY_PROBE_VOL_AVERAGE 21 /* sum of f V / sum of V over the interior cells */
#define TY_PROBE_VOL_MIN 22
#define TY_PROBE_VOL_MAX 23

/* Compile the probe QUANTITY expression (FCT syntax, FCT_PARSER.f90:113-340: + - * / ^, brackets,
 * real constants, _pi, _e, the functions inv sqr sqrt exp ln log sin cos tan sinh cosh tanh asin acos
 * atan asinh acosh atanh erf erfc abs sign step ramp, and the variables RHO, PS, V (velocity norm),
 * V_X, V_Y, V_Z, X, Y, Z, T of prb_vol_calc, PROBECALC.f90:141-233; case and spaces ignored).
 * `sign` evaluates as `inv`, as FCT_EVAL.f90:297-298 does.  Comparison and logical operators,
 * which FCT_EVAL.f90 does not evaluate, and unknown names are refused with TY_ERR_ARG and the
 * reason in ty_last_error.  *id receives the probe's index (0, 1, ... in definition order).
 * Probes belong to the context and survive ty_set_scheme and ty_set_state. */
int ty_probe_define(ty_ctx* ctx, const char* quantity, int32_t type, int32_t* id);

/* Evaluate probes ids[nprobe] on the current state at time t (prb_zone_vol.f90:34-51 after the
 * calc_varprim of write_monitors_cycle.f90:43-47): primitive fields from the FP64 conservative state,
 * the expression per interior cell on the device, reductions in a fixed order (the same bits on
 * every call).  results[nprobe] on return. */
int ty_probe_eval(ty_ctx* ctx, double t, int32_t nprobe, const int32_t* ids, double* results);

/* ty_probe_eval without device allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).  The same kernels and
 * the same results, bit for bit, with the block partials and results in a fixed 2 MiB workspace that
 * each GPU holds for the life of the process (looked up by ty_create).  Limits: at most
 * TY_PROBE_NOALLOC_MAX_CELLS interior cells and TY_PROBE_NOALLOC_MAX_PROBES probes per call, else
 * TY_ERR_ARG.  Every context on a GPU shares that GPU's workspace: calls that overlap in time on one
 * GPU (contexts driven from several threads) race and must be serialised by the caller. */
#define TY_PROBE_NOALLOC_MAX_CELLS (1 << 25)
#define TY_PROBE_NOALLOC_MAX_PROBES 64
int ty_probe_eval_noalloc(ty_ctx* ctx, double t, int32_t nprobe, const int32_t* ids, double* results);

/* per-kernel timing for benchmarks (no reference counterpart): when enabled, every
 * gradient, flux, time-step, tangent and adjoint launch is bracketed by HIP events on the
 * context's stream.  Returns the summed milliseconds and launch counts recorded since the
 * last call (ms[5], count[5]: gradient, flux, time step, tangent, adjoint) and sets the
 * enable flag. */
int ty_timing(ty_ctx* ctx, int32_t enable, double* ms, int32_t* count);

/* layout facts for the caller / harness: number of slots per cell, merged face count */
int ty_mesh_info(const ty_ctx* ctx, int32_t* ncell, int32_t* nface_dev, int32_t* nslot,
                 int32_t* nbface_dev);
/* the device cell order chosen by ty_mesh_upload (no reference counterpart; the layout a caller needs
 * to address device-ordered data): cell[i] = the caller's cell number (1-based) of the i-th device
 * cell, i = 0..ncell_int-1 */
int ty_cell_order(const ty_ctx* ctx, int32_t* cell);

/* Mesh preprocessing on the device: the faces and geometry of an element mesh, what
 * create_face_connect (CFDTOOLS/Mesh/MESHCONNECT.f90:34-163: create_facevtex.f90:15-220, the face
 * order of reorder_ustconnect MESHCONNECT.f90:511-606) and calc_meshgeom (MESHGEOM.f90:421-461)
 * compute on the host in the reference.  Three calls on a builder handle:
 *   ty_meshbuild_faces     the faces: each distinct set of element-face vertices once, numbered in
 *                          order of first appearance (element by element, local faces in the
 *                          element's order), its first element the left cell; returns the face
 *                          and internal-face counts;
 *   ty_meshbuild_boundary  the boundary faces in that order (vertices and centres), for the caller
 *                          to sort them into boundary conditions (create_ustboco,
 *                          MESHCONNECT.f90:186-278, works from the mesh file's vertex lists);
 *   ty_meshbuild_finish    with the caller's boundary order: internal faces first (in order of
 *                          appearance), then the boundary faces in that order; left cell < right
 *                          cell on internal faces (calc_ust_checkface, MESHGEOM.f90:362-410); face
 *                          geometry (calc_ust_face, MESHGEOM.f90:37-139); cell centres and volumes
 *                          (calc_ust_midcell/elemvol/cell, MESHGEOM.f90:148-352); each normal
 *                          turned to point from the left cell to the right cell (out of the domain
 *                          on the boundary).  The handle's device memory is released after it.
 * One element type per mesh.  elemvtex: cellvtex%elem(1)%elemvtex(1:ncell, 1:nv), column-major,
 * 1-based (ELEMVTEX.f90 vertex order); vertex: umesh%mesh%vertex(1:nvtex) of v3d.
 * bfacevtex(1:nbface, 1:4) column-major, 1-based, 0 where a face has fewer vertices; bcentre:
 * nbface v3d.  border[j] (j = 1..nbface): the boundary-list position (1-based) of the face that
 * becomes face nface_int + j.  Outputs of ty_meshbuild_finish in the layouts ty_mesh_desc takes:
 * facecell(1:nface, 1:2) column-major 1-based with 0 as the right cell of boundary faces (their
 * ghost cells are the caller's, init_connect_grid.f90:82-88), facevtex(1:nface, 1:4) as bfacevtex,
 * face(1:nface) st_face {normale(3), centre(3), surface}, centre(1:ncell) v3d, volume(1:ncell) of
 * the elements.  TY_ERR_ARG for a vertex number outside 1..nvtex, a face shared by more than two
 * elements, or a border that is not a permutation of 1..nbface. */
#define TY_ELEM_TRI3 13   /* ELEMVTEX.f90 element codes */
#define TY_ELEM_QUAD4 24
#define TY_ELEM_TETRA4 101
#define TY_ELEM_HEXA8 141
typedef struct ty_meshbuild ty_meshbuild;
int ty_meshbuild_create(ty_meshbuild** mb, int device);
int ty_meshbuild_destroy(ty_meshbuild* mb);
const char* ty_meshbuild_error(const ty_meshbuild* mb);
int ty_meshbuild_faces(ty_meshbuild* mb, int32_t dim, int32_t elemtype, int32_t ncell, const int32_t* elemvtex,
                       int32_t nvtex, const double* vertex, int32_t* nface, int32_t* nface_int);
int ty_meshbuild_boundary(ty_meshbuild* mb, int32_t* bfacevtex, double* bcentre);
int ty_meshbuild_finish(ty_meshbuild* mb, const int32_t* border, int32_t* facecell, int32_t* facevtex,
                        double* face, double* centre, double* volume);

#ifdef __cplusplus
}
#endif
#endif

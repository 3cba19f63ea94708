/*
 * typhon_oracle.h -- CPU restatement of Typhon's NS/Euler residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * residual path in typhon_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  It is never part of the product.
 *
 * Everything is FP64 and follows the reference's data model literally:
 *   - cells 0..ncell_int-1 are interior, ncell_int..ncell-1 are ghost cells
 *     (one per boundary face, periodic ones included);
 *   - faces carry (ic1, ic2) with ic1 < ic2 and the unit normal pointing
 *     ic1 -> ic2 (CFDTOOLS/Mesh/MESHGEOM.f90:362-410);
 *   - internal faces first (MESHCONNECT.f90:538-548), then boundary faces;
 *   - conservative variables (rho, rhoE, rho u, rho v, rho w), primitive
 *     variables (rho, p, u, v, w), stored as [var][cell].
 * Index bases are 0 here; the C-ABI of the product takes the reference's
 * 1-based arrays.
 */
#ifndef TYPHON_ORACLE_H
#define TYPHON_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* boundary condition codes: SOURCE/MODCOM/VARCOM.f90:75-93 */
#define OR_BC_SYM        10
#define OR_BC_PERIOD     11
#define OR_BC_WALL_ADIAB 20
#define OR_BC_WALL_ISOTH 21
#define OR_BC_WALL_FLUX  22
#define OR_BC_INLET_SUB  30
#define OR_BC_INLET_SUP  31
#define OR_BC_OUTLET_SUB 35
#define OR_BC_OUTLET_SUP 36

/* numerical flux codes: SOURCE/PARAM/MENU_NUM.f90:42-63 */
#define OR_SCH_RUSANOV 5
#define OR_SCH_ROE     10
#define OR_SCH_HLLE    21
#define OR_SCH_HLLEB   22
#define OR_SCH_HLLEK   23
#define OR_SCH_HLLEKB  24
#define OR_SCH_HLLC    25
#define OR_SCH_HLLCB   26
#define OR_SCH_HLLCK   27
#define OR_SCH_HLLCKB  28

/* gradient methods: MENU_NUM.f90:128-131 */
#define OR_GRAD_NONE 5
#define OR_GRAD_LSQ  10
#define OR_GRAD_LSQW 12

/* post limiters: MENU_NUM.f90:95-100 */
#define OR_POSTLIM_NONE       0
#define OR_POSTLIM_MONOTONIC0 10
#define OR_POSTLIM_MONOTONIC1 11
#define OR_POSTLIM_MONOTONIC2 12
#define OR_POSTLIM_BARTH      20
#define OR_POSTLIM_SUPERBARTH 21

/* viscous schemes: MENU_NUM.f90:120-125 */
#define OR_DIS_NOFLUX   1
#define OR_DIS_CELLAVG2 5
#define OR_DIS_CELLFULL 10

/* viscosity laws: SOURCE/EQNS/EQNS.f90:39-43 */
#define OR_VISC_SUTH   11
#define OR_VISC_DYNCST 13
#define OR_VISC_KINCST 14

typedef struct {
    int32_t type;          /* OR_BC_* */
    int32_t nface;
    const int32_t* iface;  /* 0-based face indices of this boundary */
    const int32_t* ipart;  /* periodic only: 0-based partner face per face */
    double param[8];       /* inlet_sup: ptot ttot mach dirx diry dirz;
                              outlet_sub: pstat; wall_isoth: twall;
                              wall_flux: heat flux; walls: param[4..6] = wall velocity */
} or_boco;

typedef struct {
    /* mesh (reference form) */
    int32_t dim, ncell_int, ncell, nface, nface_int;
    const int32_t *ic1, *ic2;      /* [nface], 0-based */
    const double *fnorm;           /* [nface][3] unit normal */
    const double *fcent;           /* [nface][3] face centre */
    const double *fsurf;           /* [nface] */
    const double *ccent;           /* [ncell][3] incl. ghosts */
    const double *cvol;            /* [ncell] incl. ghosts */
    int32_t nboco;
    const or_boco* boco;
    /* gas: EQNS.f90:60-70 */
    double gamma, rgas, prandtl;
    int32_t visc_law;              /* OR_VISC_* */
    double visc_dyn, visc_kin, tref, tsuth;
    int32_t viscous;               /* 0 Euler, 1 laminar NS */
    int32_t sch_dis;               /* OR_DIS_* */
    /* scheme: MENU_NUM.f90:219-231 */
    int32_t sch_hyp;
    char method;                   /* 'N' none, 'M' MUSCL, 'F' MUSCL-FAST, 'U' MUSCL-UNS */
    char limiter;                  /* 'X' none, 'M','A','V','S','K','L' */
    int32_t postlimiter;
    int32_t gradmeth;
    int32_t venkat;                /* 1: Venkatakrishnan cell limiter (no reference counterpart) */
    double venkat_k;
    /* RK tableau: rk%coef(nstage, nstage), MENU_NUM.f90:155-159, row-major coef[(i-1)*8+(j-1)] */
    int32_t nstage;
    double coef[64];
    double cfl;
} or_config;

typedef struct or_ctx or_ctx;

or_ctx* or_create(const or_config* cfg);
/* OpenMP threads the parallel loops use (1 when built without OpenMP) */
int or_threads(void);
void or_set_threads(int n);
void    or_destroy(or_ctx* o);

/* state in/out: conservative interior cells [5][ncell_int] (rho, rhoE, mx, my, mz) */
void or_set_cons(or_ctx* o, const double* cons);
void or_get_cons(const or_ctx* o, double* cons);

/* global time step (calc_ns_timestep + calc_zonetimestep, stab_cond), returns dt
   or a negative value if a cell has non-positive / NaN sound speed squared */
double or_timestep(or_ctx* o, double cfl, double dtmax);

/* one calc_rhs with global dt: residual -dt/V sum F.A on interior cells [5][ncell_int] */
void or_calc_rhs(or_ctx* o, double dt, double* res);

/* one explicit RK step with global dt (treelevel_rungekutta + update_field) */
void or_rk_step(or_ctx* o, double dt);
/* RK step with the per-cell time steps of the last or_timestep (loc_stab_cond) */
void or_rk_step_local(or_ctx* o);
/* update_field's steady residual of the last RK step (update_field.f90:42-60) */
double or_residual_norm(const or_ctx* o, double refsca0, double refsca1, double refvec);
/* calc_refcons + define_refcons: rho_ref, rhoE_ref, sqrt(rho_ref rhoE_ref) of the current state */
void or_refcons(const or_ctx* o, double* out);

/* debug views of the last calc_rhs: prim [5][ncell], grad [15][ncell], face states [5][nface], flux [5][nface] */
const double* or_prim(const or_ctx* o);
const double* or_grad(const or_ctx* o);
const double* or_face_left(const or_ctx* o);
const double* or_face_right(const or_ctx* o);
const double* or_flux(const or_ctx* o);

/* building blocks exposed for closed-form checks */
void or_flux_face(const or_config* cfg, const double* qL, const double* qR,
                  const double* n, double* f);           /* prim (rho,p,u,v,w) -> flux per unit area */
double or_limiter(char lim, double x1, double x2);       /* LIMITER.f90 scalar forms */
double or_venkat_phi(double dmax, double dmin, double d2, double eps2);
double or_viscosity(const or_config* cfg, double rho, double T); /* EQNS.f90:225-265 */

/* A15 pieces: HLL flux Jacobians of one face (row-major 5x5, dF/dU in (rho, rhoE, rho u, rho v, rho w))
   and the diagonal blocks I + dt/V dF.A/dU of the interior cells after or_calc_rhs ([ncell_int][25]) */
void or_jac_hll(const or_config* cfg, const double* qL, const double* qR, const double* n, double* jL, double* jR);
void or_jac_diag(or_ctx* o, double dt, double* diag);

/* A15, the reference's assembled implicit path (after or_calc_rhs at the same state and dt):
   block matrix over every cell incl. ghosts, diag [ncell][25], upper/lower [nface][25] row-major;
   its restarted GMRES (rhs in / solution out, block-major [ncell][5]; returns info, -1 if not
   converged); and one implicit step (U += solution on interior cells; returns the GMRES info). */
void or_bdlu_assemble(or_ctx* o, double dt, double* diag, double* upper, double* lower);
int or_bdlu_gmres(const or_ctx* o, const double* diag, const double* upper, const double* lower, double* sol,
                  int nkrylov, double maxres, int max_it);
int or_implicit_bdlu_step(or_ctx* o, double dt, int nkrylov, double maxres, int max_it);
/* residual of every cell after or_calc_rhs, ghost rows included ([5][ncell]) */
const double* or_res_all(const or_ctx* o);
/* per-cell time steps of the last or_timestep ([ncell_int]) */
const double* or_dtloc(const or_ctx* o);

/* MESHGEOM restatement (A1): geometry of faces given by vertex lists.
   fvtx [nface][4] 0-based vertex ids (-1 padded), ic1/ic2 in/out (swapped so ic1<ic2 when ic2>=0,
   ic2 = -1 on boundary faces), outputs face normal/centre/surface and cell centre/volume
   of interior cells.  Returns 0. */
int or_meshgeom(int dim, int nvtx, const double* vtx, int nface, const int32_t* fvtx,
                int32_t* ic1, int32_t* ic2, int ncell_int,
                double* fnorm, double* fcent, double* fsurf, double* ccent, double* cvol);

#ifdef __cplusplus
}
#endif
#endif

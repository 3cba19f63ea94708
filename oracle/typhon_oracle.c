/*
 * typhon_oracle.c -- CPU restatement (FP64) of Typhon's Euler/NS residual +
 * explicit RK path.  TEST INFRASTRUCTURE ONLY: the parity checker of the HIP
 * path, never linked into the product.
 *
 * OpenMP (as the reference's own !$OMP loops): loops whose iterations write
 * disjoint entries (per face or per cell) run in parallel above OR_PAR_MIN
 * iterations.  A face->cell scatter is run as a per-cell gather over the
 * cell's faces in increasing face index (a CSR built once), which adds the
 * same terms to each cell in the same order as the serial face loop: results
 * are bitwise those of the serial restatement, whatever the thread count.
 *
 * Each function cites the reference routine it restates (paths relative to the
 * reference root).  Arithmetic follows the reference's operation order where
 * that is cheap to do; the face->cell scatter runs in face order instead of the
 * reference's greedy colour order (SOURCE/MAIN/flux_to_res.f90:90-113), which
 * changes sums only at the 1e-16 relative level.
 *
 * min/max follow the semantics AMD flang gives the Fortran intrinsics
 * (MAX(a,b) = a>b ? a : b, MIN(a,b) = a<b ? a : b); this matters for the NaN
 * ratios of postlimit_barth.
 */
#include "typhon_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PAR_MIN 4096
#define FMAX(a, b) ((a) > (b) ? (a) : (b))
#define FMIN(a, b) ((a) < (b) ? (a) : (b))
#define NV 5

struct or_ctx {
    or_config c;
    int nc, ni, nf, nfi;
    double* cons;  /* [5][ncell] */
    double* prim;  /* [5][ncell] rho p u v w */
    double* grad;  /* [15][ncell]: (var*3+dir) */
    double* qL;    /* [5][nface] */
    double* qR;
    double* flux;  /* [5][nface] */
    double* res;   /* [5][ncell] */
    double* lsq;   /* [9][ncell_int] cholesky lower part, column-major 3x3 */
    int lsq_done;
    double* dtloc; /* [ncell] */
    double* ratio; /* [ncell] */
    int* cf_ptr;   /* [ncell+1] faces of each cell (interior and ghost), increasing face index */
    int* cf_idx;   /* [2 nface] */
};

#define CONS(o, v, i) ((o)->cons[(size_t)(v) * (o)->nc + (i)])
#define PRIM(o, v, i) ((o)->prim[(size_t)(v) * (o)->nc + (i)])
#define GRAD(o, v, d, i) ((o)->grad[(size_t)((v) * 3 + (d)) * (o)->nc + (i)])
#define QL(o, v, f) ((o)->qL[(size_t)(v) * (o)->nf + (f)])
#define QR(o, v, f) ((o)->qR[(size_t)(v) * (o)->nf + (f)])
#define FLX(o, v, f) ((o)->flux[(size_t)(v) * (o)->nf + (f)])
#define RES(o, v, i) ((o)->res[(size_t)(v) * (o)->nc + (i)])

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* the faces of every cell in increasing face index (counting sort over the face loop) */
static void cell_faces(or_ctx* o)
{
    if (o->cf_ptr) return;
    int nc = o->nc, nf = o->nf;
    o->cf_ptr = (int*)calloc((size_t)nc + 1, sizeof(int));
    o->cf_idx = (int*)malloc(sizeof(int) * 2 * (size_t)(nf > 0 ? nf : 1));
    for (int f = 0; f < nf; ++f) {
        o->cf_ptr[o->c.ic1[f] + 1]++;
        o->cf_ptr[o->c.ic2[f] + 1]++;
    }
    for (int i = 0; i < nc; ++i) o->cf_ptr[i + 1] += o->cf_ptr[i];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    memcpy(fill, o->cf_ptr, sizeof(int) * (size_t)nc);
    for (int f = 0; f < nf; ++f) {
        o->cf_idx[fill[o->c.ic1[f]]++] = f;
        o->cf_idx[fill[o->c.ic2[f]]++] = f;
    }
    free(fill);
}

int or_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

or_ctx* or_create(const or_config* cfg)
{
    or_ctx* o = (or_ctx*)calloc(1, sizeof(or_ctx));
    o->c = *cfg;
    o->nc = cfg->ncell;
    o->ni = cfg->ncell_int;
    o->nf = cfg->nface;
    o->nfi = cfg->nface_int;
    o->cons = (double*)calloc((size_t)NV * o->nc, sizeof(double));
    o->prim = (double*)calloc((size_t)NV * o->nc, sizeof(double));
    o->grad = (double*)calloc((size_t)15 * o->nc, sizeof(double));
    o->qL = (double*)calloc((size_t)NV * o->nf, sizeof(double));
    o->qR = (double*)calloc((size_t)NV * o->nf, sizeof(double));
    o->flux = (double*)calloc((size_t)NV * o->nf, sizeof(double));
    o->res = (double*)calloc((size_t)NV * o->nc, sizeof(double));
    o->lsq = (double*)calloc((size_t)9 * (o->ni > 0 ? o->ni : 1), sizeof(double));
    o->dtloc = (double*)calloc((size_t)o->nc, sizeof(double));
    o->ratio = (double*)calloc((size_t)o->nc, sizeof(double));
    /* ghost cells start as a copy of a neutral state so that calc_varprim on them
       stays finite (the reference computes it on ghosts too, calc_varprim_ns.f90:37) */
    for (int i = 0; i < o->nc; ++i) { CONS(o, 0, i) = 1.0; CONS(o, 1, i) = 1.0; }
    return o;
}

void or_destroy(or_ctx* o)
{
    if (!o) return;
    free(o->cons); free(o->prim); free(o->grad); free(o->qL); free(o->qR);
    free(o->flux); free(o->res); free(o->lsq); free(o->dtloc); free(o->ratio);
    free(o->cf_ptr); free(o->cf_idx);
    free(o);
}

void or_set_cons(or_ctx* o, const double* cons)
{
    for (int v = 0; v < NV; ++v)
        memcpy(&CONS(o, v, 0), cons + (size_t)v * o->ni, sizeof(double) * o->ni);
}

void or_get_cons(const or_ctx* o, double* cons)
{
    for (int v = 0; v < NV; ++v)
        memcpy(cons + (size_t)v * o->ni, &o->cons[(size_t)v * o->nc], sizeof(double) * o->ni);
}

const double* or_prim(const or_ctx* o) { return o->prim; }
const double* or_grad(const or_ctx* o) { return o->grad; }
const double* or_face_left(const or_ctx* o) { return o->qL; }
const double* or_face_right(const or_ctx* o) { return o->qR; }
const double* or_flux(const or_ctx* o) { return o->flux; }
const double* or_res_all(const or_ctx* o) { return o->res; }
const double* or_dtloc(const or_ctx* o) { return o->dtloc; }

/* ---------------------------------------------------------------------------
 * A2: calc_varprim_ns (SOURCE/EQNS/calc_varprim_ns.f90:37-57), all cells */
static void varprim(or_ctx* o)
{
    double g1 = o->c.gamma - 1.0;
#pragma omp parallel for if (o->nc > OR_PAR_MIN) schedule(static)
    for (int i = 0; i < o->nc; ++i) {
        double rho = CONS(o, 0, i);
        double u = CONS(o, 2, i) / rho, v = CONS(o, 3, i) / rho, w = CONS(o, 4, i) / rho;
        double ec = .5 * rho * (u * u + v * v + w * w);
        PRIM(o, 0, i) = rho;
        PRIM(o, 1, i) = g1 * (CONS(o, 1, i) - ec);
        PRIM(o, 2, i) = u; PRIM(o, 3, i) = v; PRIM(o, 4, i) = w;
    }
}

/* ---------------------------------------------------------------------------
 * A3: physical boundary states.  q[5] = (rho, p, u, v, w) from the interior
 * (or left face) state to the ghost (or right face) state. */

/* pi_ti_mach_dir2nspri (SOURCE/EQNS/EQNS.f90:139-160) with dir/|dir|
   (setboco_ns_inlet_sup.f90:104) */
static void bc_inlet_sup(const or_config* c, const double* prm, double* q)
{
    double g1 = c->gamma - 1.0;
    double mach = prm[2];
    double dx = prm[3], dy = prm[4], dz = prm[5];
    double dn = sqrt(dx * dx + dy * dy + dz * dz);
    dx /= dn; dy /= dn; dz /= dn;
    double fm = 1.0 / (1.0 + .5 * g1 * mach * mach);
    double ts = prm[1] * fm;
    double a = sqrt(c->gamma * c->rgas * ts);
    double p = prm[0] * pow(fm, c->gamma / g1);
    double rho = p / (c->rgas * ts);
    double ma = mach * a;
    q[0] = rho; q[1] = p; q[2] = ma * dx; q[3] = ma * dy; q[4] = ma * dz;
}

/* setboco_ns_outlet_sub.f90:57-95: nspri2pi_ti_mach_dir (EQNS.f90:197-219) of the
   interior state, then pi_ti_ps_dir2nspri (EQNS.f90:166-191) with imposed ps */
static void bc_outlet_sub(const or_config* c, const double* prm, const double* qi, double* q)
{
    double g = c->gamma, g1 = g - 1.0, r = c->rgas;
    double rho = qi[0], p = qi[1], u = qi[2], v = qi[3], w = qi[4];
    double fmv = u * u + v * v + w * w;
    double sq = sqrt(fmv);
    double dx = u / sq, dy = v / sq, dz = w / sq;
    fmv = fmv / (g * p) * rho;
    fmv = 1.0 + 0.5 * g1 * fmv;
    double pi = p * pow(fmv, g / g1);
    double ti = p / r / rho * fmv;
    double ps = prm[0];
    double fm = pow(pi / ps, g1 / g);
    if (fm < 1.0) fm = 1.0;
    double rhoo = ps / (r * ti) * fm;
    double vm = sqrt((fm - 1.0) * 2.0 / g1) * sqrt(g * ps / rhoo);
    q[0] = rhoo; q[1] = ps; q[2] = vm * dx; q[3] = vm * dy; q[4] = vm * dz;
}

/* calcboco_ust_sym.f90:46-79 (wall velocity zero): scalars copied, velocity mirrored */
static void bc_sym(const double* n, const double* qi, double* q)
{
    double vn = qi[2] * n[0] + qi[3] * n[1] + qi[4] * n[2];
    double k = 2.0 * vn;
    q[0] = qi[0]; q[1] = qi[1];
    q[2] = qi[2] - k * n[0]; q[3] = qi[3] - k * n[1]; q[4] = qi[4] - k * n[2];
}

/* setboco_ns_inlet_sub.f90:75-134 with uniform PI, TI and DIRECTION (no FCT dependence):
   ps from the interior cell (cell_state) or the left face state (face_state), dir/|dir|
   (:107), then pi_ti_ps_dir2nspri (EQNS.f90:166-191) */
static void bc_inlet_sub(const or_config* c, const double* prm, const double* qi, double* q)
{
    double g = c->gamma, g1 = g - 1.0;
    double dx = prm[3], dy = prm[4], dz = prm[5];
    double dn = sqrt(dx * dx + dy * dy + dz * dz);
    dx /= dn; dy /= dn; dz /= dn;
    double pi = prm[0], ti = prm[1], ps = qi[1];
    double fm = pow(pi / ps, g1 / g);
    if (fm < 1.0) fm = 1.0;
    double rho = ps / (c->rgas * ti) * fm;
    double vm = sqrt((fm - 1.0) * 2.0 / g1) * sqrt(g * ps / rho);
    q[0] = rho; q[1] = ps; q[2] = vm * dx; q[3] = vm * dy; q[4] = vm * dz;
}

/* setboco_ns_isoth.f90:66-84, both state modes: T_g = 2 T_w - p/(rho r), p_g = p,
   rho_g = p_g/(r T_g), u_g = 2 u_w - u (calc_wallvelocity without ALE/MRF: u_w as given) */
static void bc_wall_isoth(const or_config* c, const double* prm, const double* qi, double* q)
{
    double r = c->rgas;
    double temp = 2.0 * prm[0] - qi[1] / (qi[0] * r);
    q[1] = qi[1];
    q[0] = q[1] / (r * temp);
    for (int d = 0; d < 3; ++d) q[2 + d] = 2.0 * prm[4 + d] - qi[2 + d];
}

/* setboco_ns_flux.f90 (adiabatic wall = flux wall with zero flux, def_boco_ns.f90:44-49;
   prm[0] is the flux in the algorithm's convention, the user's WALL_FLUX negated, :73-74).
   cell_state (:79-117): d = |x_f - x_c|^2 / |(x_f - x_c).n|, conductivity at the interior
   temperature, T_g = p/(rho r) - flux d / conduct, p_g = p, rho_g = p/(r T_g), u_g = 2 u_w - u.
   face_state (:120-141): rho_R = p_L (the reference's weak condition), p_R = p_L,
   u_R = 2 u_w - u_L. */
static void bc_wall_flux(const or_ctx* o, const or_boco* b, int f, int ic, int face, const double* qi, double* q)
{
    const or_config* c = &o->c;
    double r = c->rgas;
    if (face) {
        q[0] = qi[1];
        q[1] = qi[1];
    } else {
        const double* xf = c->fcent + 3 * (size_t)f;
        const double* xc = c->ccent + 3 * (size_t)ic;
        const double* n = c->fnorm + 3 * (size_t)f;
        double dv[3] = {xf[0] - xc[0], xf[1] - xc[1], xf[2] - xc[2]};
        double d = dot3(dv, dv) / fabs(dot3(dv, n));
        double gsgmu = c->gamma / (c->gamma - 1.0);
        double cp = gsgmu * r;
        double TH = qi[1] / (r * qi[0]);
        double mu = or_viscosity(c, qi[0], TH);
        double conduct = mu * cp / c->prandtl;
        double temp = qi[1] / (qi[0] * r) - b->param[0] * d / conduct;
        q[1] = qi[1];
        q[0] = q[1] / (r * temp);
    }
    for (int dd = 0; dd < 3; ++dd) q[2 + dd] = 2.0 * b->param[4 + dd] - qi[2 + dd];
}

/* boundary state of face f from the interior/left state qi; ic is the interior cell,
   face != 0 in face_state mode */
static void bc_state(const or_ctx* o, const or_boco* b, int f, int ic, int face, const double* qi, double* q)
{
    switch (b->type) {
    case OR_BC_SYM: bc_sym(o->c.fnorm + 3 * (size_t)f, qi, q); break;
    case OR_BC_INLET_SUP: bc_inlet_sup(&o->c, b->param, q); break;
    case OR_BC_INLET_SUB: bc_inlet_sub(&o->c, b->param, qi, q); break;
    case OR_BC_OUTLET_SUB: bc_outlet_sub(&o->c, b->param, qi, q); break;
    case OR_BC_WALL_ISOTH: bc_wall_isoth(&o->c, b->param, qi, q); break;
    case OR_BC_WALL_ADIAB:
    case OR_BC_WALL_FLUX: bc_wall_flux(o, b, f, ic, face, qi, q); break;
    case OR_BC_OUTLET_SUP:
    default:
        for (int v = 0; v < NV; ++v) q[v] = qi[v];
        break;
    }
}

/* calcboco_gen mode bccon_cell_state (calcboco_gen.f90:46-88, calcboco_ust.f90:44-105,
   periodic copy calcboco_connect_per_match.f90:61-84) */
static void boco_cell_state(or_ctx* o)
{
    double qi[NV], q[NV];
    for (int ib = 0; ib < o->c.nboco; ++ib) {
        const or_boco* b = &o->c.boco[ib];
        for (int k = 0; k < b->nface; ++k) {
            int f = b->iface[k];
            int ig = o->c.ic2[f];
            int is = (b->type == OR_BC_PERIOD) ? o->c.ic1[b->ipart[k]] : o->c.ic1[f];
            for (int v = 0; v < NV; ++v) qi[v] = PRIM(o, v, is);
            if (b->type == OR_BC_PERIOD) { for (int v = 0; v < NV; ++v) q[v] = qi[v]; }
            else bc_state(o, b, f, is, 0, qi, q);
            for (int v = 0; v < NV; ++v) PRIM(o, v, ig) = q[v];
        }
    }
}

/* bccon_face_state: right face state from left face state (calc_rhs.f90:109) */
static void boco_face_state(or_ctx* o)
{
    double qi[NV], q[NV];
    for (int ib = 0; ib < o->c.nboco; ++ib) {
        const or_boco* b = &o->c.boco[ib];
        for (int k = 0; k < b->nface; ++k) {
            int f = b->iface[k];
            int is = (b->type == OR_BC_PERIOD) ? b->ipart[k] : f;
            for (int v = 0; v < NV; ++v) qi[v] = QL(o, v, is);
            if (b->type == OR_BC_PERIOD) { for (int v = 0; v < NV; ++v) q[v] = qi[v]; }
            else bc_state(o, b, f, o->c.ic1[f], 1, qi, q);
            for (int v = 0; v < NV; ++v) QR(o, v, f) = q[v];
        }
    }
}

/* bccon_cell_grad: ghost gradients.  Symmetry mirrors (calcboco_ust_sym.f90:65-77:
   vectors v-2(v.n)n, tensor S.T.S with S=I-2nn^T, VEC3D.f90:679-694); the NS
   inlet/outlet routines copy (setboco_ns_inlet_sup.f90:140-149); periodic copies.
   Isothermal wall (setboco_ns_isoth.f90:92-105): grad rho copied, grad p negated then
   mirrored (v3d_sym, VEC3D.f90:630-637), grad u copied.  Flux/adiabatic walls
   (setboco_ns_flux.f90:150-162) and the subsonic inlet (setboco_ns_inlet_sub.f90:141-154) loop
   `do ifb = ista(ib), iend(ifb)` with ifb not yet set, a loop that runs no iteration in the
   reference's runs: their ghost gradients keep the zero they start with.  The NRG viscous
   goldens (NRG/NS/NUM/FLUX/VISCOUS/CELLAVG, moving adiabatic walls, cell-average face
   gradients) are reproduced with zero ghost gradients and not with the negate/copy rule the
   loop body states. */
static void boco_cell_grad(or_ctx* o)
{
    for (int ib = 0; ib < o->c.nboco; ++ib) {
        const or_boco* b = &o->c.boco[ib];
        for (int k = 0; k < b->nface; ++k) {
            int f = b->iface[k];
            int ig = o->c.ic2[f];
            int is = (b->type == OR_BC_PERIOD) ? o->c.ic1[b->ipart[k]] : o->c.ic1[f];
            if (b->type == OR_BC_SYM) {
                const double* n = o->c.fnorm + 3 * (size_t)f;
                for (int v = 0; v < 2; ++v) {
                    double g[3] = {GRAD(o, v, 0, is), GRAD(o, v, 1, is), GRAD(o, v, 2, is)};
                    double k2 = 2.0 * (g[0] * n[0] + g[1] * n[1] + g[2] * n[2]);
                    for (int d = 0; d < 3; ++d) GRAD(o, v, d, ig) = g[d] - k2 * n[d];
                }
                double S[3][3], T[3][3], M[3][3];
                S[0][0] = 1.0 - 2.0 * n[0] * n[0]; S[1][1] = 1.0 - 2.0 * n[1] * n[1];
                S[2][2] = 1.0 - 2.0 * n[2] * n[2];
                S[0][1] = -2.0 * n[0] * n[1]; S[0][2] = -2.0 * n[0] * n[2];
                S[1][0] = S[0][1]; S[1][2] = -2.0 * n[1] * n[2];
                S[2][0] = S[0][2]; S[2][1] = S[1][2];
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) T[i][j] = GRAD(o, 2 + i, j, is);
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                    double s = 0; for (int l = 0; l < 3; ++l) s += T[i][l] * S[l][j]; M[i][j] = s; }
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                    double s = 0; for (int l = 0; l < 3; ++l) s += S[i][l] * M[l][j];
                    GRAD(o, 2 + i, j, ig) = s; }
            } else {
                for (int v = 0; v < 15; ++v) o->grad[(size_t)v * o->nc + ig] = o->grad[(size_t)v * o->nc + is];
                if (b->type == OR_BC_WALL_ISOTH) {
                    const double* n = o->c.fnorm + 3 * (size_t)f;
                    double g[3] = {-GRAD(o, 1, 0, is), -GRAD(o, 1, 1, is), -GRAD(o, 1, 2, is)};
                    double k2 = 2.0 * (g[0] * n[0] + g[1] * n[1] + g[2] * n[2]);
                    for (int d = 0; d < 3; ++d) GRAD(o, 1, d, ig) = g[d] - k2 * n[d];
                } else if (b->type == OR_BC_WALL_FLUX || b->type == OR_BC_WALL_ADIAB ||
                           b->type == OR_BC_INLET_SUB) {
                    for (int v = 0; v < 15; ++v) o->grad[(size_t)v * o->nc + ig] = 0.0;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * A4: precalc_grad_lsq (SOURCE/MGRID/precalc_grad_lsq.f90:60-155) and
 * cholesky_decomp (SOURCE/MATH/MATRIX.f90:42-70) */
/* weighted LSQ: the matrix uses d/|d|^2 (precalc_grad_lsq.f90:70-73) while the
   right-hand side uses d/|d|^4 (calc_gradient.f90:94-101, sqrabs(dvec)**2), i.e. both
   are the weight 1/|d|^4 of d d^T and dq d.  rhs selects the right-hand-side form. */
static void face_dcg(const or_ctx* o, int f, double* d, int rhs)
{
    const double* c1 = o->c.ccent + 3 * (size_t)o->c.ic1[f];
    const double* c2 = o->c.ccent + 3 * (size_t)o->c.ic2[f];
    d[0] = c2[0] - c1[0]; d[1] = c2[1] - c1[1]; d[2] = c2[2] - c1[2];
    if (o->c.gradmeth == OR_GRAD_LSQW) {
        double s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (rhs) s = s * s;
        d[0] /= s; d[1] /= s; d[2] /= s;
    }
}

static void cholesky_decomp(double* a /* 3x3 column-major a[i+3j] */)
{
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            double sum = a[i + 3 * j];
            for (int k = i - 1; k >= 0; --k) sum = sum - a[i + 3 * k] * a[j + 3 * k];
            if (i == j) a[i + 3 * i] = sqrt(sum);
            else a[j + 3 * i] = sum / a[i + 3 * i];
        }
}

static void precalc_lsq(or_ctx* o)
{
    memset(o->lsq, 0, sizeof(double) * 9 * (size_t)o->ni);
    for (int f = 0; f < o->nf; ++f) {
        double d[3]; face_dcg(o, f, d, 0);
        double m[9];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i + 3 * j] = d[i] * d[j];
        m[0] = d[0] * d[0]; m[4] = d[1] * d[1]; m[8] = d[2] * d[2];
        int c1 = o->c.ic1[f];
        for (int k = 0; k < 9; ++k) o->lsq[9 * (size_t)c1 + k] += m[k];
        if (f < o->nfi) {
            int c2 = o->c.ic2[f];
            for (int k = 0; k < 9; ++k) o->lsq[9 * (size_t)c2 + k] += m[k];
        }
    }
    if (o->c.dim == 2)
        for (int i = 0; i < o->ni; ++i) o->lsq[9 * (size_t)i + 8] += 1.0;
    for (int i = 0; i < o->ni; ++i) cholesky_decomp(o->lsq + 9 * (size_t)i);
    o->lsq_done = 1;
}

/* A5: calc_gradient (SOURCE/MGRID/calc_gradient.f90:75-192) + cholesky_solve
   (MATRIX.f90:79-111).  Order of variables: rho, p, then velocity components
   (tensor rows mat(i,1:3) = grad u_i). */
static void calc_gradient(or_ctx* o)
{
    if (!o->lsq_done) precalc_lsq(o);
    cell_faces(o);
    int ni = o->ni;
    /* right-hand side sum_f (q_2 - q_1) d_f over the cell's faces (an interior cell is ic2 only
       of two-sided faces), then the Cholesky solve */
#pragma omp parallel for if (ni > OR_PAR_MIN) schedule(static)
    for (int i = 0; i < ni; ++i) {
        double rhs[NV * 3] = {0.0};
        for (int p = o->cf_ptr[i]; p < o->cf_ptr[i + 1]; ++p) {
            int f = o->cf_idx[p];
            double d[3]; face_dcg(o, f, d, 1);
            int c1 = o->c.ic1[f], c2 = o->c.ic2[f];
            for (int v = 0; v < NV; ++v) {
                double ds = PRIM(o, v, c2) - PRIM(o, v, c1);
                double* r = rhs + v * 3;
                r[0] += ds * d[0]; r[1] += ds * d[1]; r[2] += ds * d[2];
            }
        }
        const double* a = o->lsq + 9 * (size_t)i;
        for (int v = 0; v < NV; ++v) {
            double* b = rhs + v * 3;
            for (int r = 0; r < 3; ++r) {
                for (int j = 0; j < r; ++j) b[r] = b[r] - a[r + 3 * j] * b[j];
                b[r] = b[r] / a[r + 3 * r];
            }
            for (int r = 2; r >= 0; --r) {
                for (int j = r + 1; j < 3; ++j) b[r] = b[r] - a[j + 3 * r] * b[j];
                b[r] = b[r] / a[r + 3 * r];
            }
            for (int d = 0; d < 3; ++d) GRAD(o, v, d, i) = b[d];
        }
    }
}

/* ---------------------------------------------------------------------------
 * A6: limiters (SOURCE/MGRID/LIMITER.f90:72-339) */
static double sgn1(double x) { return copysign(1.0, x); }

double or_limiter(char lim, double x1, double x2)
{
    double r = 0.0;
    switch (lim) {
    case 'M': { /* minmod :72 */
        double s = .5 * (sgn1(x1) + sgn1(x2));
        r = s * FMIN(fabs(x1), fabs(x2));
        break; }
    case 'A': { /* albada :122 */
        double s = x1 * x2;
        r = (s <= 0.0) ? 0.0 : s * (x1 + x2) / (x1 * x1 + x2 * x2);
        break; }
    case 'V': { /* vleer :175 */
        double s = x1 * x2;
        r = (s <= 0.0) ? 0.0 : 2.0 * s / (x1 + x2);
        break; }
    case 'S': /* superbee :227 */
        if (x1 > 0.0 && x2 > 0.0) r = FMAX(FMIN(2.0 * x1, x2), FMIN(2.0 * x2, x1));
        if (x1 < 0.0 && x2 < 0.0) r = FMIN(FMAX(2.0 * x1, x2), FMAX(2.0 * x2, x1));
        break;
    case 'K': /* kim3 :277 */
        if (x1 > 0.0 && x2 > 0.0) { r = (2.0 * x1 + x2) / 3.0; r = FMIN(FMIN(2.0 * x1, 2.0 * x2), r); }
        if (x1 < 0.0 && x2 < 0.0) { r = (2.0 * x1 + x2) / 3.0; r = FMAX(FMAX(2.0 * x1, 2.0 * x2), r); }
        break;
    case 'L': { /* lim03 :339 */
        double eta = (x1 * x1 + x2 * x2) / 0.01 / 0.01;
        if (eta < 1.0) r = (2.0 * x2 + x1) / 3.0;
        else if (x1 > 0.0 && x2 > 0.0) { r = (2.0 * x2 + x1) / 3.0; r = FMIN(FMIN(1.6 * x2, 2.0 * x1), r); }
        else if (x1 < 0.0 && x2 < 0.0) { r = (2.0 * x2 + x1) / 3.0; r = FMAX(FMAX(1.6 * x2, 2.0 * x1), r); }
        else if (x1 < 0.0 && x2 > 0.0) { r = (2.0 * x2 + x1) / 3.0; r = FMAX(0.0, FMIN(-0.5 * x1, r)); }
        else if (x1 > 0.0 && x2 < 0.0) { r = (2.0 * x2 + x1) / 3.0; r = FMIN(0.0, FMAX(-0.5 * x1, r)); }
        break; }
    default: r = 0.0;
    }
    return r;
}

/* slope correction shared by tvdgradstr (MUSCL, :35-58) and tvdgraduns (MUSCL-UNS):
   returns the increment added to the projected slope s. */
static double muscl_corr(char method, char lim, double s, double lr)
{
    if (lim == 'X') return lr - s;
    if (method == 'M') {
        if (lim == 'K') return or_limiter(lim, lr, 2.0 * s - lr) - s;
        return or_limiter(lim, 2.0 * s - lr, lr) - s;
    }
    /* MUSCL-UNS (tvdgraduns_scal.f90:35-57) */
    if (lim == 'K') return or_limiter(lim, lr, s) - s;
    return or_limiter(lim, s, lr) - s;
}

/* MUSCL-FAST slope (tvdgradfst_scal.f90:35-60): limited slope along LR */
static double fast_slope(char lim, double s, double lr)
{
    if (lim == 'X') return lr;
    if (lim == 'K') return or_limiter(lim, lr, 2.0 * s - lr);
    return or_limiter(lim, 2.0 * s - lr, lr);
}

/* calc_hres_states (SOURCE/MAIN/calc_hres_states.f90:40-96) -> hres_ns_muscl
   (hres_ns_muscl.f90:51-143), hres_ns_musclfast (:51-144), hres_ns_muscluns, or
   distrib_field for first order; all faces, ghost cells on the right side. */
static void hres_states(or_ctx* o)
{
    const or_config* c = &o->c;
#pragma omp parallel for if (o->nf > OR_PAR_MIN) schedule(static)
    for (int f = 0; f < o->nf; ++f) {
        int cl = c->ic1[f], cr = c->ic2[f];
        if (c->method == 'N') {
            for (int v = 0; v < NV; ++v) { QL(o, v, f) = PRIM(o, v, cl); QR(o, v, f) = PRIM(o, v, cr); }
            continue;
        }
        const double* xl = c->ccent + 3 * (size_t)cl;
        const double* xr = c->ccent + 3 * (size_t)cr;
        const double* xf = c->fcent + 3 * (size_t)f;
        double u[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
        double LF[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
        double RF[3] = {xf[0] - xr[0], xf[1] - xr[1], xf[2] - xr[2]};
        if (c->method == 'F') {
            double kl = sqrt(LF[0] * LF[0] + LF[1] * LF[1] + LF[2] * LF[2]);
            double kr = -sqrt(RF[0] * RF[0] + RF[1] * RF[1] + RF[2] * RF[2]);
            double dlr = kl - kr;
            u[0] /= dlr; u[1] /= dlr; u[2] /= dlr;
            for (int v = 0; v < NV; ++v) {
                double lr = (PRIM(o, v, cr) - PRIM(o, v, cl)) / dlr;
                double gl[3] = {GRAD(o, v, 0, cl), GRAD(o, v, 1, cl), GRAD(o, v, 2, cl)};
                double gr[3] = {GRAD(o, v, 0, cr), GRAD(o, v, 1, cr), GRAD(o, v, 2, cr)};
                double sl = (c->limiter == 'X') ? 0.0 : dot3(gl, u);
                double sr = (c->limiter == 'X') ? 0.0 : dot3(gr, u);
                QL(o, v, f) = PRIM(o, v, cl) + fast_slope(c->limiter, sl, lr) * kl;
                QR(o, v, f) = PRIM(o, v, cr) + fast_slope(c->limiter, sr, lr) * kr;
            }
        } else { /* 'M' or 'U' */
            double dlr = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            u[0] /= dlr; u[1] /= dlr; u[2] /= dlr;
            for (int v = 0; v < NV; ++v) {
                double lr = (PRIM(o, v, cr) - PRIM(o, v, cl)) / dlr;
                for (int side = 0; side < 2; ++side) {
                    int ci = side ? cr : cl;
                    const double* cf = side ? RF : LF;
                    double g[3] = {GRAD(o, v, 0, ci), GRAD(o, v, 1, ci), GRAD(o, v, 2, ci)};
                    double s = dot3(g, u);
                    double ds = muscl_corr(c->method, c->limiter, s, lr);
                    g[0] = g[0] + ds * u[0]; g[1] = g[1] + ds * u[1]; g[2] = g[2] + ds * u[2];
                    double dq = dot3(g, cf);
                    if (side) QR(o, v, f) = PRIM(o, v, cr) + dq;
                    else QL(o, v, f) = PRIM(o, v, cl) + dq;
                }
            }
        }
    }
}

/* A7: postlimit_barth (SOURCE/MGRID/postlimit_barth.f90:50-223), face order,
   separate ratios for rho, p and the velocity vector (all three components). */
static void postlimit_barth(or_ctx* o, double klim)
{
    const or_config* c = &o->c;
    static const int groups[3][2] = {{0, 1}, {1, 2}, {2, 5}};
    for (int g = 0; g < 3; ++g) {
        for (int i = 0; i < o->nc; ++i) o->ratio[i] = 1.0;
        for (int f = 0; f < o->nf; ++f) {
            int cl = c->ic1[f], cr = c->ic2[f];
            for (int v = groups[g][0]; v < groups[g][1]; ++v) {
                double lr = PRIM(o, v, cr) - PRIM(o, v, cl);
                double l = lr / (QL(o, v, f) - PRIM(o, v, cl));
                l = FMAX(klim * l, 0.0);
                o->ratio[cl] = FMIN(o->ratio[cl], l);
            }
            for (int v = groups[g][0]; v < groups[g][1]; ++v) {
                double lr = PRIM(o, v, cr) - PRIM(o, v, cl);
                double l = -lr / (QR(o, v, f) - PRIM(o, v, cr));
                l = FMAX(klim * l, 0.0);
                o->ratio[cr] = FMIN(o->ratio[cr], l);
            }
        }
        for (int f = 0; f < o->nf; ++f) {
            int cl = c->ic1[f], cr = c->ic2[f];
            for (int v = groups[g][0]; v < groups[g][1]; ++v) {
                QL(o, v, f) = PRIM(o, v, cl) + o->ratio[cl] * (QL(o, v, f) - PRIM(o, v, cl));
                QR(o, v, f) = PRIM(o, v, cr) + o->ratio[cr] * (QR(o, v, f) - PRIM(o, v, cr));
            }
        }
    }
}

/* A7: monotonic0/1/2 (SOURCE/MGRID/LIMITER.f90:427-477) of the face values x1 (left) and x2
   (right) between the cell values a (left) and b (right); postlimit_monotonic
   (SOURCE/MGRID/postlimit_monotonic.f90:40-110) applies them on every face, ghost cells on the
   right, to rho, p and each velocity component. */
static void monotonic(int code, double* x1, double* x2, double a, double b)
{
    double l = *x1, r = *x2;
    if (code == OR_POSTLIM_MONOTONIC0) {
        if (a < b) { l = FMAX(a, FMIN(l, b)); r = FMIN(b, FMAX(r, a)); }
        else       { l = FMIN(a, FMAX(l, b)); r = FMAX(b, FMIN(r, a)); }
    } else if (code == OR_POSTLIM_MONOTONIC1) {
        double midx = 0.5 * (l + r);
        if (a < b) { l = FMAX(a, FMIN(FMIN(l, midx), b)); r = FMIN(b, FMAX(FMAX(r, midx), a)); }
        else       { l = FMIN(a, FMAX(FMAX(l, midx), b)); r = FMAX(b, FMIN(FMIN(r, midx), a)); }
    } else {
        double mid = 0.5 * (a + b);
        if (a < b) { l = FMAX(a, FMIN(l, mid)); r = FMIN(b, FMAX(r, mid)); }
        else       { l = FMIN(a, FMAX(l, mid)); r = FMAX(b, FMIN(r, mid)); }
    }
    *x1 = l; *x2 = r;
}

static void postlimit_monotonic(or_ctx* o, int code)
{
    const or_config* c = &o->c;
    for (int f = 0; f < o->nf; ++f) {
        int cl = c->ic1[f], cr = c->ic2[f];
        for (int v = 0; v < NV; ++v)
            monotonic(code, &QL(o, v, f), &QR(o, v, f), PRIM(o, v, cl), PRIM(o, v, cr));
    }
}

/* A8: Venkatakrishnan limiter (no reference counterpart; Venkatakrishnan, J. Comput.
   Phys. 118 (1995) 120-130, eq. 4.4 smooth min with eps^2 = (K h)^3, h = V^(1/dim)).
   Written as phi = [(dm^2+eps2) + 2 d2 dm] / [dm^2 + 2 d2^2 + dm d2 + eps2]
   where dm = dmax (d2 > 0) or dmin (d2 < 0); phi = 1 when d2 == 0. */
double or_venkat_phi(double dmax, double dmin, double d2, double eps2)
{
    if (d2 == 0.0) return 1.0;
    double dm = (d2 > 0.0) ? dmax : dmin;
    double num = (dm * dm + eps2) + 2.0 * d2 * dm;
    double den = dm * dm + 2.0 * d2 * d2 + dm * d2 + eps2;
    if (den <= 0.0) return 1.0;
    double phi = num / den;
    return FMIN(1.0, FMAX(0.0, phi));
}

/* Venkatakrishnan cell limiter applied to the LSQ gradient, then a plain
   unlimited MUSCL face extrapolation q_f = q_c + phi grad(q).(x_f - x_c). */
static void venkat_states(or_ctx* o)
{
    const or_config* c = &o->c;
    int nc = o->nc;
    double* qmax = (double*)malloc(sizeof(double) * NV * (size_t)nc);
    double* qmin = (double*)malloc(sizeof(double) * NV * (size_t)nc);
    double* phi = (double*)malloc(sizeof(double) * NV * (size_t)nc);
    cell_faces(o);
    double k3 = c->venkat_k * c->venkat_k * c->venkat_k;
    /* per cell, over its faces in increasing face index: the neighbour extrema, then the minimum
       of the face limiter values.  Ghost cells keep qmax = qmin = q and phi = 1 (they are the
       second cell of one-sided faces only, which update nothing). */
#pragma omp parallel for if (nc > OR_PAR_MIN) schedule(static)
    for (int i = 0; i < nc; ++i) {
        for (int v = 0; v < NV; ++v) {
            qmax[(size_t)v * nc + i] = PRIM(o, v, i);
            qmin[(size_t)v * nc + i] = PRIM(o, v, i);
            phi[(size_t)v * nc + i] = 1.0;
        }
        if (i >= o->ni) continue;
        for (int p = o->cf_ptr[i]; p < o->cf_ptr[i + 1]; ++p) {
            int f = o->cf_idx[p];
            int cl = c->ic1[f], cr = c->ic2[f];
            if (cl != i && f >= o->nfi) continue;
            int other = (cl == i) ? cr : cl;
            for (int v = 0; v < NV; ++v) {
                size_t ii = (size_t)v * nc + i;
                double b = PRIM(o, v, other);
                qmax[ii] = FMAX(qmax[ii], b); qmin[ii] = FMIN(qmin[ii], b);
            }
        }
        const double* xc = c->ccent + 3 * (size_t)i;
        double h = (c->dim == 2) ? sqrt(c->cvol[i]) : cbrt(c->cvol[i]);
        double eps2 = k3 * h * h * h;
        for (int p = o->cf_ptr[i]; p < o->cf_ptr[i + 1]; ++p) {
            int f = o->cf_idx[p];
            if (c->ic1[f] != i && f >= o->nfi) continue;
            const double* xf = c->fcent + 3 * (size_t)f;
            double r[3] = {xf[0] - xc[0], xf[1] - xc[1], xf[2] - xc[2]};
            for (int v = 0; v < NV; ++v) {
                double g[3] = {GRAD(o, v, 0, i), GRAD(o, v, 1, i), GRAD(o, v, 2, i)};
                double d2 = dot3(g, r);
                size_t ii = (size_t)v * nc + i;
                double pv = or_venkat_phi(qmax[ii] - PRIM(o, v, i), qmin[ii] - PRIM(o, v, i), d2, eps2);
                phi[ii] = FMIN(phi[ii], pv);
            }
        }
    }
#pragma omp parallel for if (o->nf > OR_PAR_MIN) schedule(static)
    for (int f = 0; f < o->nf; ++f) {
        int cl = c->ic1[f], cr = c->ic2[f];
        for (int side = 0; side < 2; ++side) {
            int ci = side ? cr : cl;
            const double* xc = c->ccent + 3 * (size_t)ci;
            const double* xf = c->fcent + 3 * (size_t)f;
            double r[3] = {xf[0] - xc[0], xf[1] - xc[1], xf[2] - xc[2]};
            for (int v = 0; v < NV; ++v) {
                double g[3] = {GRAD(o, v, 0, ci), GRAD(o, v, 1, ci), GRAD(o, v, 2, ci)};
                double q = PRIM(o, v, ci) + phi[(size_t)v * nc + ci] * dot3(g, r);
                if (side) QR(o, v, f) = q; else QL(o, v, f) = q;
            }
        }
    }
    free(qmax); free(qmin); free(phi);
}

/* ---------------------------------------------------------------------------
 * A9: numerical fluxes per unit area.  q = (rho, p, u, v, w). */
static void roe_state(double g, const double* qL, const double* qR, double* rho, double* vel, double* p)
{
    /* calc_roe_states (SOURCE/EQNS/calc_roe_states.f90:35-53) */
    double gig1 = g / (g - 1.0);
    double t1 = sqrt(qR[0] / qL[0]);
    *rho = qL[0] * t1;
    double kl = 1.0 / (1.0 + t1);
    double kr = t1 * kl;
    for (int d = 0; d < 3; ++d) vel[d] = kl * qL[2 + d] + kr * qR[2 + d];
    double hl = gig1 * qL[1] / qL[0] + .5 * (qL[2] * qL[2] + qL[3] * qL[3] + qL[4] * qL[4]);
    double hr = gig1 * qR[1] / qR[0] + .5 * (qR[2] * qR[2] + qR[3] * qR[3] + qR[4] * qR[4]);
    double h = kl * hl + kr * hr;
    *p = (h - .5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])) * (*rho) / gig1;
}

#define OR_PI 3.141592653589793238462643 /* PIcst, CFDTOOLS/Math/MATH.f90:15 */

/* Wave speeds of the HLL family (calc_flux_hlle.f90:64-123, calc_flux_hllc.f90:64-125): Einfeldt's
 * bounds with beta = 1 (HLLE, HLLC) or sqrt((g-1)/2g) (the B variants), HLLE's bounded by zero; the
 * kinetic variants (HLLEK/HLLEKB/HLLCK/HLLCKB) replace the side bounds by the mean speeds of the
 * side's half Maxwellian, c h(M) with c = a / sqrt(g/2), M = vn/c -+ sqrt(g/2) beta. */
static void hll_speeds(int sch, double g, const double* qL, const double* qR, const double* n, double vnl, double vnr,
                       double* sl, double* sr)
{
    int kin = (sch == OR_SCH_HLLEK || sch == OR_SCH_HLLEKB || sch == OR_SCH_HLLCK || sch == OR_SCH_HLLCKB);
    double beta = (sch == OR_SCH_HLLEB || sch == OR_SCH_HLLEKB || sch == OR_SCH_HLLCB || sch == OR_SCH_HLLCKB)
                  ? sqrt((g - 1.0) / (2 * g)) : 1.0;
    double al = sqrt(g * qL[1] / qL[0]), ar = sqrt(g * qR[1] / qR[0]);
    double rrho, rv[3], rp; roe_state(g, qL, qR, &rrho, rv, &rp);
    double vm = rv[0] * n[0] + rv[1] * n[1] + rv[2] * n[2];
    double am = sqrt(g * rp / rrho);
    if (kin) {
        double f1 = sqrt(0.5 * g), f2 = 0.5 / sqrt(OR_PI);
        double cl = al / f1, cr = ar / f1;
        double ml = vnl / cl - f1 * beta, mr = vnr / cr + f1 * beta;
        double kinwl = 0.5 * (1.0 - erf(ml)), kinwr = 0.5 * (1.0 + erf(mr));
        double kindl = -f2 * exp(-ml * ml), kindr = +f2 * exp(-mr * mr);
        *sl = FMIN(cl * (ml * kinwl + kindl), vm - am);
        *sr = FMAX(cr * (mr * kinwr + kindr), vm + am);
    } else if (sch == OR_SCH_HLLE || sch == OR_SCH_HLLEB) {
        *sl = FMIN(FMIN(0.0, vnl - beta * al), vm - am);
        *sr = FMAX(FMAX(0.0, vnr + beta * ar), vm + am);
    } else {
        *sl = FMIN(vnl - beta * al, vm - am);
        *sr = FMAX(vnr + beta * ar, vm + am);
    }
}

void or_flux_face(const or_config* c, const double* qL, const double* qR, const double* n, double* F)
{
    double g = c->gamma, ig1 = 1.0 / (g - 1.0);
    double vnl = qL[2] * n[0] + qL[3] * n[1] + qL[4] * n[2];
    double vnr = qR[2] * n[0] + qR[3] * n[1] + qR[4] * n[2];
    double rel = ig1 * qL[1] + .5 * qL[0] * (qL[2] * qL[2] + qL[3] * qL[3] + qL[4] * qL[4]);
    double rer = ig1 * qR[1] + .5 * qR[0] * (qR[2] * qR[2] + qR[3] * qR[3] + qR[4] * qR[4]);
    switch (c->sch_hyp) {
    case OR_SCH_RUSANOV: { /* calc_flux_rusanov.f90:64-83 */
        double al = sqrt(g * qL[1] / qL[0]), ar = sqrt(g * qR[1] / qR[0]);
        double ray = FMAX(fabs(vnl) + al, fabs(vnr) + ar);
        F[0] = .5 * ((vnl + ray) * qL[0] + (vnr - ray) * qR[0]);
        F[1] = .5 * ((vnl + ray) * rel + (vnr - ray) * rer + vnl * qL[1] + vnr * qR[1]);
        for (int d = 0; d < 3; ++d)
            F[2 + d] = .5 * (((vnl + ray) * qL[0]) * qL[2 + d] + ((vnr - ray) * qR[0]) * qR[2 + d]
                             + (qL[1] + qR[1]) * n[d]);
        return; }
    case OR_SCH_HLLE: case OR_SCH_HLLEB: case OR_SCH_HLLEK: case OR_SCH_HLLEKB: { /* calc_flux_hlle.f90:109-143 */
        double sl, sr;
        hll_speeds(c->sch_hyp, g, qL, qR, n, vnl, vnr, &sl, &sr);
        double iks = 1.0 / (sr - sl);
        double kl = sr * iks, kr = -sl * iks, ku = sl * sr * iks;
        F[0] = (kl * vnl - ku) * qL[0] + (kr * vnr + ku) * qR[0];
        F[1] = (kl * vnl - ku) * rel + (kr * vnr + ku) * rer + (kl * vnl * qL[1] + kr * vnr * qR[1]);
        for (int d = 0; d < 3; ++d)
            F[2 + d] = ((kl * vnl - ku) * qL[0]) * qL[2 + d] + ((kr * vnr + ku) * qR[0]) * qR[2 + d]
                       + (kl * qL[1] + kr * qR[1]) * n[d];
        return; }
    case OR_SCH_ROE: { /* textbook Roe with Harten entropy fix (no reference counterpart:
                          calc_flux_inviscid.f90:77-78 stops on ROE) */
        double rrho, rv[3], rp; roe_state(g, qL, qR, &rrho, rv, &rp);
        double H = (g / (g - 1.0)) * rp / rrho + .5 * (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
        double a = sqrt(g * rp / rrho);
        double vn = rv[0] * n[0] + rv[1] * n[1] + rv[2] * n[2];
        double drho = qR[0] - qL[0], dp = qR[1] - qL[1], dvn = vnr - vnl;
        double dv[3] = {qR[2] - qL[2], qR[3] - qL[3], qR[4] - qL[4]};
        double l1 = fabs(vn - a), l2 = fabs(vn), l3 = fabs(vn + a);
        double del = 0.1 * a;
        if (l1 < del) l1 = .5 * (l1 * l1 / del + del);
        if (l3 < del) l3 = .5 * (l3 * l3 / del + del);
        double w1 = l1 * (dp - rrho * a * dvn) / (2 * a * a);
        double w2 = l2 * (drho - dp / (a * a));
        double w3 = l3 * (dp + rrho * a * dvn) / (2 * a * a);
        double w4 = l2 * rrho;
        double dvt[3] = {dv[0] - dvn * n[0], dv[1] - dvn * n[1], dv[2] - dvn * n[2]};
        double FL[5], FR[5];
        FL[0] = qL[0] * vnl; FR[0] = qR[0] * vnr;
        FL[1] = vnl * (rel + qL[1]); FR[1] = vnr * (rer + qR[1]);
        for (int d = 0; d < 3; ++d) { FL[2 + d] = qL[0] * vnl * qL[2 + d] + qL[1] * n[d];
                                      FR[2 + d] = qR[0] * vnr * qR[2 + d] + qR[1] * n[d]; }
        double D[5];
        D[0] = w1 + w2 + w3;
        D[1] = w1 * (H - vn * a) + w2 * .5 * (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2])
               + w3 * (H + vn * a) + w4 * (rv[0] * dvt[0] + rv[1] * dvt[1] + rv[2] * dvt[2]);
        for (int d = 0; d < 3; ++d)
            D[2 + d] = w1 * (rv[d] - a * n[d]) + w2 * rv[d] + w3 * (rv[d] + a * n[d]) + w4 * dvt[d];
        for (int k = 0; k < 5; ++k) F[k] = .5 * (FL[k] + FR[k] - D[k]);
        return; }
    case OR_SCH_HLLC: case OR_SCH_HLLCB: case OR_SCH_HLLCK: case OR_SCH_HLLCKB: default: { /* calc_flux_hllc.f90:61-171 */
        double sl, sr;
        hll_speeds(c->sch_hyp, g, qL, qR, n, vnl, vnr, &sl, &sr);
        if (sl >= 0.0) {
            F[0] = vnl * qL[0]; F[1] = vnl * (rel + qL[1]);
            for (int d = 0; d < 3; ++d) F[2 + d] = (vnl * qL[0]) * qL[2 + d] + qL[1] * n[d];
        } else if (sr <= 0.0) {
            F[0] = vnr * qR[0]; F[1] = vnr * (rer + qR[1]);
            for (int d = 0; d < 3; ++d) F[2 + d] = (vnr * qR[0]) * qR[2 + d] + qR[1] * n[d];
        } else {
            double rqL = qL[0] * (sl - vnl), rqR = qR[0] * (sr - vnr);
            double sst = (rqR * vnr - rqL * vnl - qR[1] + qL[1]) / (rqR - rqL);
            double iks, rst, pst, rvst[3], rest;
            if (sst >= 0.0) {
                iks = 1.0 / (sl - sst); rst = rqL * iks; pst = qL[1] - rqL * (vnl - sst);
                for (int d = 0; d < 3; ++d) rvst[d] = iks * ((rqL * qL[2 + d]) + (pst - qL[1]) * n[d]);
                rest = iks * (rel * (sl - vnl) - qL[1] * vnl + pst * sst);
            } else {
                iks = 1.0 / (sr - sst); rst = rqR * iks; pst = qR[1] - rqR * (vnr - sst);
                for (int d = 0; d < 3; ++d) rvst[d] = iks * ((rqR * qR[2 + d]) + (pst - qR[1]) * n[d]);
                rest = iks * (rer * (sr - vnr) - qR[1] * vnr + pst * sst);
            }
            F[0] = sst * rst; F[1] = sst * (rest + pst);
            for (int d = 0; d < 3; ++d) F[2 + d] = (sst * rvst[d]) + (pst * n[d]);
        }
        return; }
    }
}

/* ---------------------------------------------------------------------------
 * A10: viscous fluxes, calc_flux_viscous_cell (SOURCE/EQNS/calc_flux_viscous_cell.f90:75-158)
 * with interp_facegradient_vect (SOURCE/MESH/interpface_gradient_vect.f90:59-81),
 * interp_facegradn_scal (SOURCE/MESH/interpface_gradn_scal.f90:59-91) and
 * calc_viscosity (SOURCE/EQNS/EQNS.f90:225-265).  Cell states and cell gradients
 * are used on both sides (integration_ns_ust.f90:136-159); ghost cells carry the
 * cell_state / cell_grad boundary values. */
double or_viscosity(const or_config* c, double rho, double T)
{
    switch (c->visc_law) {
    case OR_VISC_SUTH: return c->visc_dyn * pow(T / c->tref, 1.5) * (c->tref + c->tsuth) / (T + c->tsuth);
    case OR_VISC_KINCST: return c->visc_kin * rho;
    case OR_VISC_DYNCST: default: return c->visc_dyn;
    }
}

static void viscous_fluxes(or_ctx* o)
{
    const or_config* c = &o->c;
    double r = c->rgas;
    double cp = c->gamma * r / (c->gamma - 1);
    for (int f = 0; f < o->nf; ++f) {
        int cl = c->ic1[f], cr = c->ic2[f];
        const double* xl = c->ccent + 3 * (size_t)cl;
        const double* xr = c->ccent + 3 * (size_t)cr;
        const double* xf = c->fcent + 3 * (size_t)f;
        const double* n = c->fnorm + 3 * (size_t)f;
        double a[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
        double b[3] = {xf[0] - xr[0], xf[1] - xr[1], xf[2] - xr[2]};
        double dHL = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double dHR = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        double id = 1.0 / (dHL + dHR);
        dHL = id * dHL; dHR = id * dHR;
        double vLR[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
        double rhoL = PRIM(o, 0, cl), rhoR = PRIM(o, 0, cr);
        double pL = PRIM(o, 1, cl), pR = PRIM(o, 1, cr);
        double vL[3] = {PRIM(o, 2, cl), PRIM(o, 3, cl), PRIM(o, 4, cl)};
        double vR[3] = {PRIM(o, 2, cr), PRIM(o, 3, cr), PRIM(o, 4, cr)};
        double velH[3];
        for (int d = 0; d < 3; ++d) velH[d] = dHR * vL[d] + dHL * vR[d];
        double TL = pL / (rhoL * r), TR = pR / (rhoR * r);
        double TH = dHR * TL + dHL * TR;
        double rhoH = dHR * rhoL + dHL * rhoR;
        double gTL[3], gTR[3];
        double kl1 = 1.0 / (rhoL * r), kl2 = pL / (rhoL * rhoL * r);
        double kr1 = 1.0 / (rhoR * r), kr2 = pR / (rhoR * rhoR * r);
        for (int d = 0; d < 3; ++d) {
            gTL[d] = kl1 * GRAD(o, 1, d, cl) - kl2 * GRAD(o, 0, d, cl);
            gTR[d] = kr1 * GRAD(o, 1, d, cr) - kr2 * GRAD(o, 0, d, cr);
        }
        double mu = or_viscosity(c, rhoH, TH);
        /* face velocity gradient G(i,j) = d u_i / d x_j */
        double G[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) G[i][j] = dHL * GRAD(o, 2 + i, j, cr) + dHR * GRAD(o, 2 + i, j, cl);
        double dLR2 = vLR[0] * vLR[0] + vLR[1] * vLR[1] + vLR[2] * vLR[2];
        if (c->sch_dis == OR_DIS_CELLFULL) {
            double kv[3] = {(1.0 / dLR2) * vLR[0], (1.0 / dLR2) * vLR[1], (1.0 / dLR2) * vLR[2]};
            for (int i = 0; i < 3; ++i) {
                double gv = G[i][0] * vLR[0] + G[i][1] * vLR[1] + G[i][2] * vLR[2];
                double fa = vR[i] - vL[i] - gv;
                for (int j = 0; j < 3; ++j) G[i][j] = G[i][j] + fa * kv[j];
            }
        }
        double gTH;
        double gavg[3] = {dHL * gTR[0] + dHR * gTL[0], dHL * gTR[1] + dHR * gTL[1], dHL * gTR[2] + dHR * gTL[2]};
        if (c->sch_dis == OR_DIS_CELLFULL) {
            double pscal = (vLR[0] * n[0] + vLR[1] * n[1] + vLR[2] * n[2]) / dLR2;
            double fcomp = pscal * (TR - TL);
            double vi[3] = {n[0] - pscal * vLR[0], n[1] - pscal * vLR[1], n[2] - pscal * vLR[2]};
            gTH = 1.0 * fcomp + dot3(gavg, vi);
        } else {
            gTH = dot3(gavg, n);
        }
        double S[3][3];
        double tr = G[0][0] + G[1][1] + G[2][2];
        double addv = -(2.0 / 3.0) * tr;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) S[i][j] = G[i][j] + G[j][i];
        for (int i = 0; i < 3; ++i) S[i][i] = S[i][i] + addv;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) S[i][j] = mu * S[i][j];
        double sn[3];
        for (int i = 0; i < 3; ++i) sn[i] = S[i][0] * n[0] + S[i][1] * n[1] + S[i][2] * n[2];
        double svn = dot3(sn, velH);
        for (int d = 0; d < 3; ++d) FLX(o, 2 + d, f) = FLX(o, 2 + d, f) - sn[d];
        FLX(o, 1, f) = FLX(o, 1, f) - svn;
        double conduct = mu * cp / c->prandtl;
        FLX(o, 1, f) = FLX(o, 1, f) - conduct * gTH;
    }
}

/* ns_bocoflux (SOURCE/EQNS/ns_bocoflux.f90:47-111) on flux and adiabatic walls: the
   conductive flux is recomputed from the cell states and gradients of the face's two cells,
   as the viscous flux computed it, and added back together with the imposed flux (the
   boundary field holds the flux set by the last state-mode call, setboco_ns_flux.f90:104,129). */
static void ns_bocoflux(or_ctx* o)
{
    const or_config* c = &o->c;
    double r = c->rgas;
    double cp = c->gamma * r / (c->gamma - 1);
    for (int ib = 0; ib < c->nboco; ++ib) {
        const or_boco* b = &c->boco[ib];
        if (b->type != OR_BC_WALL_FLUX && b->type != OR_BC_WALL_ADIAB) continue;
        for (int k = 0; k < b->nface; ++k) {
            int f = b->iface[k];
            int cl = c->ic1[f], cr = c->ic2[f];
            const double* xl = c->ccent + 3 * (size_t)cl;
            const double* xr = c->ccent + 3 * (size_t)cr;
            const double* xf = c->fcent + 3 * (size_t)f;
            const double* n = c->fnorm + 3 * (size_t)f;
            double a[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
            double bb[3] = {xf[0] - xr[0], xf[1] - xr[1], xf[2] - xr[2]};
            double dHL = sqrt(dot3(a, a)), dHR = sqrt(dot3(bb, bb));
            double id = 1.0 / (dHL + dHR);
            dHL = id * dHL; dHR = id * dHR;
            double vLR[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
            double rhoL = PRIM(o, 0, cl), rhoR = PRIM(o, 0, cr);
            double TL = PRIM(o, 1, cl) / (rhoL * r), TR = PRIM(o, 1, cr) / (rhoR * r);
            double TH = dHR * TL + dHL * TR, rhoH = dHR * rhoL + dHL * rhoR;
            double gTL[3], gTR[3];
            for (int d = 0; d < 3; ++d) {
                gTL[d] = 1.0 / (rhoL * r) * GRAD(o, 1, d, cl) - PRIM(o, 1, cl) / (rhoL * rhoL * r) * GRAD(o, 0, d, cl);
                gTR[d] = 1.0 / (rhoR * r) * GRAD(o, 1, d, cr) - PRIM(o, 1, cr) / (rhoR * rhoR * r) * GRAD(o, 0, d, cr);
            }
            double mu = or_viscosity(c, rhoH, TH);
            double gavg[3] = {dHL * gTR[0] + dHR * gTL[0], dHL * gTR[1] + dHR * gTL[1], dHL * gTR[2] + dHR * gTL[2]};
            double gTH;
            if (c->sch_dis == OR_DIS_CELLFULL) {
                double dLR2 = dot3(vLR, vLR);
                double pscal = dot3(vLR, n) / dLR2;
                double vi[3] = {n[0] - pscal * vLR[0], n[1] - pscal * vLR[1], n[2] - pscal * vLR[2]};
                gTH = 1.0 * (pscal * (TR - TL)) + dot3(gavg, vi);
            } else {
                gTH = dot3(gavg, n);
            }
            double conduct = mu * cp / c->prandtl;
            FLX(o, 1, f) = FLX(o, 1, f) + conduct * gTH + b->param[0];
        }
    }
}

/* ---------------------------------------------------------------------------
 * calc_rhs (SOURCE/MAIN/calc_rhs.f90:48-145) for one grid with global dt. */
static void calc_rhs_dt(or_ctx* o, double dt, const double* dtl);
static void calc_rhs_internal(or_ctx* o, double dt) { calc_rhs_dt(o, dt, NULL); }

/* dtl: per-cell time steps (loc_stab_cond) or NULL (the global dt) */
static void calc_rhs_dt(or_ctx* o, double dt, const double* dtl)
{
    const or_config* c = &o->c;
    varprim(o);
    int second = (c->method != 'N');
    int needgrad = second || c->viscous;
    if (needgrad) {
        boco_cell_state(o);
        calc_gradient(o);
        boco_cell_grad(o);
    }
    if (second && c->venkat) {
        venkat_states(o);
    } else {
        hres_states(o);
        if (c->postlimiter == OR_POSTLIM_BARTH) postlimit_barth(o, 0.5);
        else if (c->postlimiter == OR_POSTLIM_SUPERBARTH) postlimit_barth(o, 1.0);
        else if (c->postlimiter >= OR_POSTLIM_MONOTONIC0 && c->postlimiter <= OR_POSTLIM_MONOTONIC2)
            postlimit_monotonic(o, c->postlimiter);
    }
    boco_face_state(o);
    /* integration_ns_ust (EQNS/integration_ns_ust.f90:87-192): inviscid fluxes */
#pragma omp parallel for if (o->nf > OR_PAR_MIN) schedule(static)
    for (int f = 0; f < o->nf; ++f) {
        double ql[NV], qr[NV], F[NV];
        for (int v = 0; v < NV; ++v) { ql[v] = QL(o, v, f); qr[v] = QR(o, v, f); }
        or_flux_face(c, ql, qr, c->fnorm + 3 * (size_t)f, F);
        for (int v = 0; v < NV; ++v) FLX(o, v, f) = F[v];
    }
    if (c->viscous) {
        viscous_fluxes(o);
        ns_bocoflux(o);
    }
    /* flux_to_res (SOURCE/MAIN/flux_to_res.f90:47-142) */
#pragma omp parallel for if (o->nf > OR_PAR_MIN) schedule(static)
    for (int f = 0; f < o->nf; ++f)
        for (int v = 0; v < NV; ++v) FLX(o, v, f) = c->fsurf[f] * FLX(o, v, f);
    /* R_c = sum over the cell's faces in face order of -F (c = ic1) or +F (c = ic2) */
    cell_faces(o);
#pragma omp parallel for if (o->nc > OR_PAR_MIN) schedule(static)
    for (int i = 0; i < o->nc; ++i) {
        double r[NV] = {0.0};
        for (int p = o->cf_ptr[i]; p < o->cf_ptr[i + 1]; ++p) {
            int f = o->cf_idx[p];
            if (c->ic1[f] == i)
                for (int v = 0; v < NV; ++v) r[v] = r[v] - FLX(o, v, f);
            else
                for (int v = 0; v < NV; ++v) r[v] = r[v] + FLX(o, v, f);
        }
        for (int v = 0; v < NV; ++v) RES(o, v, i) = r[v];
    }
#pragma omp parallel for if (o->ni > OR_PAR_MIN) schedule(static)
    for (int i = 0; i < o->ni; ++i) {
        double coef = (dtl ? dtl[i] : dt) / c->cvol[i];
        for (int v = 0; v < NV; ++v) RES(o, v, i) = coef * RES(o, v, i);
    }
}

void or_calc_rhs(or_ctx* o, double dt, double* res)
{
    calc_rhs_internal(o, dt);
    if (res)
        for (int v = 0; v < NV; ++v) memcpy(res + (size_t)v * o->ni, &RES(o, v, 0), sizeof(double) * o->ni);
}

/* A13: calc_ns_timestep (SOURCE/EQNS/calc_ns_timestep.f90:52-102) and
   calc_zonetimestep (SOURCE/ZONE/calc_zonetimestep.f90:57-126, stab_cond). */
double or_timestep(or_ctx* o, double cfl, double dtmax)
{
    const or_config* c = &o->c;
    int ni = o->ni;
    double gg1 = c->gamma * (c->gamma - 1.0);
    for (int i = 0; i < ni; ++i) o->dtloc[i] = 0.0;
    for (int f = 0; f < o->nf; ++f) {
        int c1 = c->ic1[f], c2 = c->ic2[f];
        o->dtloc[c1] = o->dtloc[c1] + c->fsurf[f];
        if (c2 < ni) o->dtloc[c2] = o->dtloc[c2] + c->fsurf[f];
    }
    double dt = HUGE_VAL;
    int bad = 0;
    for (int i = 0; i < ni; ++i) {
        double d = cfl * 2.0 * c->cvol[i] / o->dtloc[i];
        double irho = 1.0 / CONS(o, 0, i);
        double rv2 = CONS(o, 2, i) * CONS(o, 2, i) + CONS(o, 3, i) * CONS(o, 3, i) + CONS(o, 4, i) * CONS(o, 4, i);
        double a2 = (CONS(o, 1, i) - .5 * rv2 * irho) * gg1 * irho;
        if (!(a2 > 0.0)) bad++;
        d = d / (sqrt(rv2) * irho + sqrt(a2));
        o->dtloc[i] = d;
        dt = FMIN(dt, d);
    }
    if (bad) return -1.0;
    dt = FMIN(dt, dtmax);
    return dt;
}

static void rk_step_dt(or_ctx* o, double dt, const double* dtl);

/* A14: treelevel_rungekutta (SOURCE/MAIN/treelevel_rungekutta.f90:50-113) and
   update_field (SOURCE/ZONE/update_field.f90:38) with global dt. */
void or_rk_step(or_ctx* o, double dt) { rk_step_dt(o, dt, NULL); }

/* the same with the per-cell time steps of the last or_timestep (loc_stab_cond,
   calc_zonetimestep.f90:64-66: dtloc is not reset to the minimum) */
void or_rk_step_local(or_ctx* o) { rk_step_dt(o, 0.0, o->dtloc); }

/* A14: update_field's steady residual (SOURCE/ZONE/update_field.f90:42-60) of the last RK step:
   sum over interior cells of |residu| / refsca for rho and rhoE, plus |residu momentum| (the
   vector's norm) / refvec, with residu the step's increment. */
double or_residual_norm(const or_ctx* o, double refsca0, double refsca1, double refvec)
{
    int ni = o->ni;
    double r = 0.0, s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < ni; ++i) s0 += fabs(RES(o, 0, i));
    r = r + s0 / refsca0;
    for (int i = 0; i < ni; ++i) s1 += fabs(RES(o, 1, i));
    r = r + s1 / refsca1;
    for (int i = 0; i < ni; ++i) {
        double m = sqrt(RES(o, 2, i) * RES(o, 2, i) + RES(o, 3, i) * RES(o, 3, i) + RES(o, 4, i) * RES(o, 4, i));
        r = r + m / refvec;
    }
    return r;
}

/* calc_refcons (SOURCE/ZONE/calc_refcons.f90:20-117) with define_refcons
   (SOURCE/PARAM/MENU_SOLVER.f90:262-276) for the NS solver: volume averages of rho and rhoE over
   the interior cells, and the momentum reference sqrt(rho_ref rhoE_ref).  out[3]. */
void or_refcons(const or_ctx* o, double* out)
{
    const or_config* c = &o->c;
    double v = 0.0, s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < o->ni; ++i) {
        v += c->cvol[i];
        s0 += CONS(o, 0, i) * c->cvol[i];
        s1 += CONS(o, 1, i) * c->cvol[i];
    }
    out[0] = s0 / v;
    out[1] = s1 / v;
    out[2] = sqrt(out[0] * out[1]);
}

static void rk_step_dt(or_ctx* o, double dt, const double* dtl)
{
    const or_config* c = &o->c;
    int ns = c->nstage, nc = o->nc;
    size_t sz = (size_t)NV * nc;
    double* u0 = (double*)malloc(sizeof(double) * sz);
    double* rhs = (double*)malloc(sizeof(double) * sz * (size_t)ns);
    memcpy(u0, o->cons, sizeof(double) * sz);
    for (int s = 1; s <= ns; ++s) {
        if (s >= 2) {
            memcpy(o->cons, u0, sizeof(double) * sz);
            for (int j = 1; j <= s - 1; ++j) {
                double a = c->coef[(s - 2) * 8 + (j - 1)];
                const double* r = rhs + (size_t)(j - 1) * sz;
#pragma omp parallel for if (sz > OR_PAR_MIN) schedule(static)
                for (size_t k = 0; k < sz; ++k) o->cons[k] = o->cons[k] + a * r[k];
            }
        }
        calc_rhs_dt(o, dt, dtl);
        memcpy(rhs + (size_t)(s - 1) * sz, o->res, sizeof(double) * sz);
    }
    /* residu = coef(n,1) R1 + sum coef(n,j) Rj ; U = U0 + residu */
    double b1 = c->coef[(ns - 1) * 8 + 0];
#pragma omp parallel for if (sz > OR_PAR_MIN) schedule(static)
    for (size_t k = 0; k < sz; ++k) o->res[k] = rhs[k] * b1;
    for (int j = 2; j <= ns; ++j) {
        double b = c->coef[(ns - 1) * 8 + (j - 1)];
        const double* r = rhs + (size_t)(j - 1) * sz;
#pragma omp parallel for if (sz > OR_PAR_MIN) schedule(static)
        for (size_t k = 0; k < sz; ++k) o->res[k] = o->res[k] + b * r[k];
    }
#pragma omp parallel for if (sz > OR_PAR_MIN) schedule(static)
    for (size_t k = 0; k < sz; ++k) o->cons[k] = u0[k] + o->res[k];
    free(u0); free(rhs);
}

/* ---------------------------------------------------------------------------
 * A1: MESHGEOM restatement (CFDTOOLS/Mesh/MESHGEOM.f90:37-461). */
int or_meshgeom(int dim, int nvtx, const double* vtx, int nface, const int32_t* fvtx,
                int32_t* ic1, int32_t* ic2, int ncell_int,
                double* fnorm, double* fcent, double* fsurf, double* ccent, double* cvol)
{
    (void)nvtx;
    /* calc_ust_face :37-139 */
    for (int f = 0; f < nface; ++f) {
        const int32_t* fv = fvtx + 4 * (size_t)f;
        int ns = 0; while (ns < 4 && fv[ns] >= 0) ns++;
        const double* v1 = vtx + 3 * (size_t)fv[0];
        const double* v2 = vtx + 3 * (size_t)fv[1];
        double nrm[3], surf, cen[3];
        if (dim == 2) {
            double e[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
            nrm[0] = 0.0 * e[2] - 1.0 * e[1]; nrm[1] = 1.0 * e[0] - 0.0 * e[2]; nrm[2] = 0.0 * e[1] - 0.0 * e[0];
            surf = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            for (int d = 0; d < 3; ++d) cen[d] = .5 * (v1[d] + v2[d]);
        } else {
            const double* v3 = vtx + 3 * (size_t)fv[2];
            double a[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
            double b[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
            double n1[3] = {.5 * (a[1] * b[2] - a[2] * b[1]), .5 * (a[2] * b[0] - a[0] * b[2]), .5 * (a[0] * b[1] - a[1] * b[0])};
            if (ns == 3) {
                for (int d = 0; d < 3; ++d) { nrm[d] = n1[d]; cen[d] = (v1[d] + v2[d] + v3[d]) / 3.0; }
                surf = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            } else {
                const double* v4 = vtx + 3 * (size_t)fv[3];
                double s1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
                double c4[3] = {v4[0] - v1[0], v4[1] - v1[1], v4[2] - v1[2]};
                double n2[3] = {.5 * (b[1] * c4[2] - b[2] * c4[1]), .5 * (b[2] * c4[0] - b[0] * c4[2]), .5 * (b[0] * c4[1] - b[1] * c4[0])};
                double s2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
                for (int d = 0; d < 3; ++d) {
                    double cg1 = (v1[d] + v2[d] + v3[d]) / 3.0, cg2 = (v1[d] + v3[d] + v4[d]) / 3.0;
                    nrm[d] = n1[d] + n2[d];
                    cen[d] = (s1 * cg1 + s2 * cg2) / (s1 + s2);
                }
                surf = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            }
        }
        for (int d = 0; d < 3; ++d) { fnorm[3 * (size_t)f + d] = nrm[d] / surf; fcent[3 * (size_t)f + d] = cen[d]; }
        fsurf[f] = surf;
    }
    /* calc_ust_midcell :148-197 */
    double* mid = (double*)calloc(3 * (size_t)ncell_int, sizeof(double));
    int* nb = (int*)calloc((size_t)ncell_int, sizeof(int));
    for (int f = 0; f < nface; ++f) {
        int a = ic1[f], b = ic2[f];
        for (int d = 0; d < 3; ++d) mid[3 * (size_t)a + d] += fcent[3 * (size_t)f + d];
        nb[a]++;
        if (b >= 0 && b < ncell_int) { for (int d = 0; d < 3; ++d) mid[3 * (size_t)b + d] += fcent[3 * (size_t)f + d]; nb[b]++; }
    }
    for (int i = 0; i < ncell_int; ++i) for (int d = 0; d < 3; ++d) mid[3 * (size_t)i + d] /= (double)nb[i];
    /* calc_ust_elemvol :215-277 and calc_ust_cell :286-331 */
    double kcg = (dim == 2) ? 1. / 3.0 : 1. / 4.0;
    double kvol = (dim == 2) ? 1. / 2.0 : 1. / 3.0;
    for (int i = 0; i < ncell_int; ++i) { cvol[i] = 0.0; ccent[3 * (size_t)i] = ccent[3 * (size_t)i + 1] = ccent[3 * (size_t)i + 2] = 0.0; }
    for (int f = 0; f < nface; ++f) {
        for (int side = 0; side < 2; ++side) {
            int ci = side ? ic2[f] : ic1[f];
            if (side && !(ci >= 0 && ci < ncell_int)) continue;
            const double* fc = fcent + 3 * (size_t)f;
            const double* fn = fnorm + 3 * (size_t)f;
            double kh[3] = {fc[0] - mid[3 * (size_t)ci], fc[1] - mid[3 * (size_t)ci + 1], fc[2] - mid[3 * (size_t)ci + 2]};
            double h = fabs(kh[0] * fn[0] + kh[1] * fn[1] + kh[2] * fn[2]);
            double vol = kvol * h * fsurf[f];
            for (int d = 0; d < 3; ++d) ccent[3 * (size_t)ci + d] += vol * (fc[d] - kcg * kh[d]);
            cvol[ci] += vol;
        }
    }
    for (int i = 0; i < ncell_int; ++i) for (int d = 0; d < 3; ++d) ccent[3 * (size_t)i + d] /= cvol[i];
    /* calc_ust_checkface :362-410 */
    for (int f = 0; f < nface; ++f) {
        if (ic2[f] >= 0 && ic1[f] > ic2[f]) { int t = ic1[f]; ic1[f] = ic2[f]; ic2[f] = t; }
        const double* a = ccent + 3 * (size_t)ic1[f];
        double v12[3];
        if (ic2[f] >= 0) { const double* b = ccent + 3 * (size_t)ic2[f]; for (int d = 0; d < 3; ++d) v12[d] = b[d] - a[d]; }
        else { const double* b = fcent + 3 * (size_t)f; for (int d = 0; d < 3; ++d) v12[d] = b[d] - a[d]; }
        double* n = fnorm + 3 * (size_t)f;
        if (v12[0] * n[0] + v12[1] * n[1] + v12[2] * n[2] < 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
    }
    free(mid); free(nb);
    return 0;
}

/* ---------------------------------------------------------------------------
 * A15: flux Jacobians and the diagonal blocks of the implicit operator.
 * calc_jacflux (SOURCE/EQNS/calc_jacflux.f90:49-77): Euler flux Jacobian A(n) in
 * (rho, rhoE, rho u, rho v, rho w) at a primitive state; calc_jac_hll
 * (calc_jac_hll.f90:46-57) with the flux's wave speeds clipped at zero
 * (calc_flux_hllc.f90:185-189, calc_flux_hlle.f90:158-160).  Blocks are row-major m[i*5+j] = dF_i/dU_j. */
static void jacflux(double g, const double* q, const double* n, double vn, double* m)
{
    double g1 = g - 1.0, ig1 = 1.0 / g1;
    double v2 = q[2] * q[2] + q[3] * q[3] + q[4] * q[4];
    double en = ig1 * q[1] / q[0] + .5 * v2;
    const double* u = q + 2;
    for (int k = 0; k < 25; ++k) m[k] = 0.0;
    m[0 * 5 + 0] = 0.0; m[0 * 5 + 1] = 0.0;
    for (int j = 0; j < 3; ++j) m[0 * 5 + 2 + j] = n[j];
    m[1 * 5 + 0] = vn * (g1 * v2 - g * en);
    m[1 * 5 + 1] = vn * g;
    for (int j = 0; j < 3; ++j) m[1 * 5 + 2 + j] = (g * en - .5 * g1 * v2) * n[j] - g1 * vn * u[j];
    for (int i = 0; i < 3; ++i) {
        m[(2 + i) * 5 + 0] = .5 * g1 * v2 * n[i] - vn * u[i];
        m[(2 + i) * 5 + 1] = g1 * n[i];
        for (int j = 0; j < 3; ++j) m[(2 + i) * 5 + 2 + j] = u[i] * n[j] - (g1 * n[i]) * u[j];
        m[(2 + i) * 5 + 2 + i] += vn;
    }
}

void or_jac_hll(const or_config* c, const double* qL, const double* qR, const double* n, double* jL, double* jR)
{
    double g = c->gamma;
    double vnl = qL[2] * n[0] + qL[3] * n[1] + qL[4] * n[2];
    double vnr = qR[2] * n[0] + qR[3] * n[1] + qR[4] * n[2];
    /* the flux's own wave speeds (HLLC's for the fluxes outside the HLL family, which the
       reference does not pair with HLL Jacobians) */
    int sch = (c->sch_hyp >= OR_SCH_HLLE && c->sch_hyp <= OR_SCH_HLLCKB) ? c->sch_hyp : OR_SCH_HLLC;
    double sl, sr;
    hll_speeds(sch, g, qL, qR, n, vnl, vnr, &sl, &sr);
    sl = FMIN(sl, 0.0);
    sr = FMAX(sr, 0.0);
    jacflux(g, qL, n, vnl, jL);
    jacflux(g, qR, n, vnr, jR);
    for (int i = 0; i < 5; ++i) { jL[i * 5 + i] = jL[i * 5 + i] - sl; jR[i * 5 + i] = jR[i * 5 + i] - sr; }
    double kl = sr / (sr - sl), kr = -sl / (sr - sl);
    for (int k = 0; k < 25; ++k) { jL[k] = kl * jL[k]; jR[k] = kr * jR[k]; }
}

/* Diagonal blocks of the implicit operator I + dt/V dF.A/dU of interior cells
 * (init_implicit_bdlu.f90:37-49, build_implicit_bdlu.f90:44-83) from the face states
 * of the last calc_rhs; viscous diagonal of calc_flux_viscous_cell.f90:149-158.
 * diag: [ncell_int][25] row-major. */
void or_jac_diag(or_ctx* o, double dt, double* diag)
{
    const or_config* c = &o->c;
    int ni = o->ni;
    for (int i = 0; i < ni; ++i)
        for (int k = 0; k < 25; ++k) diag[(size_t)25 * i + k] = (k % 6 == 0) ? 1.0 : 0.0;
    for (int f = 0; f < o->nf; ++f) {
        double ql[NV], qr[NV], jL[25], jR[25];
        for (int v = 0; v < NV; ++v) { ql[v] = QL(o, v, f); qr[v] = QR(o, v, f); }
        const double* n = c->fnorm + 3 * (size_t)f;
        or_jac_hll(c, ql, qr, n, jL, jR);
        if (c->viscous) {
            int cl = c->ic1[f], cr = c->ic2[f];
            const double* xl = c->ccent + 3 * (size_t)cl;
            const double* xr = c->ccent + 3 * (size_t)cr;
            const double* xf = c->fcent + 3 * (size_t)f;
            double a[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
            double b[3] = {xf[0] - xr[0], xf[1] - xr[1], xf[2] - xr[2]};
            double dHL = sqrt(dot3(a, a)), dHR = sqrt(dot3(b, b));
            double id0 = 1.0 / (dHL + dHR);
            dHL = id0 * dHL; dHR = id0 * dHR;
            double rL = PRIM(o, 0, cl), rR = PRIM(o, 0, cr);
            double TL = PRIM(o, 1, cl) / (rL * c->rgas), TR = PRIM(o, 1, cr) / (rR * c->rgas);
            double TH = dHR * TL + dHL * TR, rhoH = dHR * rL + dHL * rR;
            double mu = or_viscosity(c, rhoH, TH);
            double d[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
            double id = mu / (dHR * rL + dHL * rR) / sqrt(dot3(d, d));
            for (int i = 0; i < 5; ++i) { jR[i * 5 + i] = jR[i * 5 + i] - id; jL[i * 5 + i] = jL[i * 5 + i] + id; }
        }
        double s = c->fsurf[f];
        int c1 = c->ic1[f], c2 = c->ic2[f];
        double k1 = dt / c->cvol[c1];
        for (int k = 0; k < 25; ++k) diag[(size_t)25 * c1 + k] += (s * jL[k]) * k1;
        if (f < o->nfi) {
            double k2 = dt / c->cvol[c2];
            for (int k = 0; k < 25; ++k) diag[(size_t)25 * c2 + k] -= (s * jR[k]) * k2;
        }
    }
}

/* ---------------------------------------------------------------------------
 * A15: the assembled implicit path.  Face Jacobians at the face states (calc_jac_hll.f90:46-57,
 * calc_jacflux.f90:49-77), the viscous diagonal (calc_flux_viscous_cell.f90:149-158), the boundary
 * Jacobians of ns_bocojacobian.f90:44-92 (SYMMETRY: jL += jR dQR/dQL with the mirror on the
 * momentum block, jR = 0; SUPERSONIC_OUTLET: jL += jR, jR = 0; others unchanged), both times the
 * face area (flux_to_res.f90:75-81), assembled as build_implicit_bdlu.f90:44-83 over every cell
 * including ghosts (init_implicit_bdlu.f90:37-49: interior diagonals start at I; a ghost row is I). */
void or_bdlu_assemble(or_ctx* o, double dt, double* diag, double* upper, double* lower)
{
    const or_config* c = &o->c;
    for (int i = 0; i < o->ni; ++i)
        for (int k = 0; k < 25; ++k) diag[(size_t)25 * i + k] = (k % 6 == 0) ? 1.0 : 0.0;
    for (int f = 0; f < o->nf; ++f) {
        double ql[NV], qr[NV], jL[25], jR[25];
        for (int v = 0; v < NV; ++v) { ql[v] = QL(o, v, f); qr[v] = QR(o, v, f); }
        const double* n = c->fnorm + 3 * (size_t)f;
        or_jac_hll(c, ql, qr, n, jL, jR);
        int cl = c->ic1[f], cr = c->ic2[f];
        if (c->viscous) {
            const double* xl = c->ccent + 3 * (size_t)cl;
            const double* xr = c->ccent + 3 * (size_t)cr;
            const double* xf = c->fcent + 3 * (size_t)f;
            double a[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
            double b[3] = {xf[0] - xr[0], xf[1] - xr[1], xf[2] - xr[2]};
            double dHL = sqrt(dot3(a, a)), dHR = sqrt(dot3(b, b));
            double id0 = 1.0 / (dHL + dHR);
            dHL = id0 * dHL; dHR = id0 * dHR;
            double rL = PRIM(o, 0, cl), rR = PRIM(o, 0, cr);
            double TL = PRIM(o, 1, cl) / (rL * c->rgas), TR = PRIM(o, 1, cr) / (rR * c->rgas);
            double TH = dHR * TL + dHL * TR, rhoH = dHR * rL + dHL * rR;
            double mu = or_viscosity(c, rhoH, TH);
            double d[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
            double id = mu / (dHR * rL + dHL * rR) / sqrt(dot3(d, d));
            for (int i = 0; i < 5; ++i) { jR[i * 5 + i] = jR[i * 5 + i] - id; jL[i * 5 + i] = jL[i * 5 + i] + id; }
        }
        if (f >= o->nfi) {
            int bt = -1;
            for (int ib = 0; ib < c->nboco && bt < 0; ++ib)
                for (int k = 0; k < c->boco[ib].nface; ++k)
                    if (c->boco[ib].iface[k] == f) { bt = c->boco[ib].type; break; }
            if (bt == OR_BC_SYM || bt == OR_BC_OUTLET_SUP) {
                double S[25];
                for (int k = 0; k < 25; ++k) S[k] = (k % 6 == 0) ? 1.0 : 0.0;
                if (bt == OR_BC_SYM)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) S[(2 + i) * 5 + 2 + j] = S[(2 + i) * 5 + 2 + j] - 2.0 * (n[i] * n[j]);
                for (int i = 0; i < 5; ++i)
                    for (int j = 0; j < 5; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < 5; ++k) s += jR[i * 5 + k] * S[k * 5 + j];
                        jL[i * 5 + j] = jL[i * 5 + j] + s;
                    }
                for (int k = 0; k < 25; ++k) jR[k] = 0.0;
            }
        }
        double s = c->fsurf[f];
        for (int k = 0; k < 25; ++k) { jL[k] = s * jL[k]; jR[k] = s * jR[k]; }
        for (int k = 0; k < 25; ++k) {
            diag[(size_t)25 * cl + k] = diag[(size_t)25 * cl + k] + jL[k] * dt / c->cvol[cl];
            upper[(size_t)25 * f + k] = jR[k] * dt / c->cvol[cl];
        }
        if (f < o->nfi) {
            for (int k = 0; k < 25; ++k) {
                diag[(size_t)25 * cr + k] = diag[(size_t)25 * cr + k] - jR[k] * dt / c->cvol[cr];
                lower[(size_t)25 * f + k] = -jL[k] * dt / c->cvol[cr];
            }
        } else {
            for (int k = 0; k < 25; ++k) {
                diag[(size_t)25 * cr + k] = (k % 6 == 0) ? 1.0 : 0.0;
                lower[(size_t)25 * f + k] = 0.0;
            }
        }
    }
}

/* bdlu_yeqax (SPMAT_BDLU.f90:139-163): y = A x, block-major vectors [ncell][5] */
static void bdlu_yeqax(const or_ctx* o, const double* diag, const double* upper, const double* lower,
                       const double* x, double* y)
{
    for (int i = 0; i < o->nc; ++i)
        for (int r = 0; r < 5; ++r) {
            double s = 0.0;
            for (int k = 0; k < 5; ++k) s += diag[(size_t)25 * i + r * 5 + k] * x[(size_t)5 * i + k];
            y[(size_t)5 * i + r] = s;
        }
    for (int f = 0; f < o->nf; ++f) {
        int i1 = o->c.ic1[f], i2 = o->c.ic2[f];
        for (int r = 0; r < 5; ++r) {
            double s = 0.0;
            for (int k = 0; k < 5; ++k) s += lower[(size_t)25 * f + r * 5 + k] * x[(size_t)5 * i1 + k];
            y[(size_t)5 * i2 + r] = y[(size_t)5 * i2 + r] + s;
        }
        for (int r = 0; r < 5; ++r) {
            double s = 0.0;
            for (int k = 0; k < 5; ++k) s += upper[(size_t)25 * f + r * 5 + k] * x[(size_t)5 * i2 + k];
            y[(size_t)5 * i1 + r] = y[(size_t)5 * i1 + r] + s;
        }
    }
}

static double vdot(const double* a, const double* b, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* bdlu_gmres (SOURCE/MATH/bdlu_gmres.f90:55-212): restarted GMRES(nkrylov), modified Gram-Schmidt,
   Givens rotations, no preconditioner; x0 = rhs / diag(ib,ib); restarts while
   ||rhs - A x|| / ||rhs|| >= ||x0||_1 maxres (the reference's mixed scales) and nit <= 2 max_it.
   sol holds the right-hand side on entry and the solution on exit; returns info (nit - 1, or -1). */
int or_bdlu_gmres(const or_ctx* o, const double* diag, const double* upper, const double* lower, double* sol,
                  int nkrylov, double maxres, int max_it)
{
    const size_t dim = (size_t)5 * o->nc;
    const int m = nkrylov;
    double* p1 = (double*)malloc(sizeof(double) * dim);
    double* r1 = (double*)malloc(sizeof(double) * dim);
    double* v1 = (double*)malloc(sizeof(double) * dim * (size_t)m);
    double* w1 = (double*)malloc(sizeof(double) * dim);
    double* hh = (double*)calloc((size_t)(m + 1) * m, sizeof(double)); /* hh(i,j) = hh[(j-1)*(m+1) + i-1] */
    double* ss = (double*)calloc((size_t)m + 1, sizeof(double));
    double* qc = (double*)calloc((size_t)m, sizeof(double));
    double* qs = (double*)calloc((size_t)m, sizeof(double));
    double* yy = (double*)calloc((size_t)m, sizeof(double));
#define HH(i, j) hh[(size_t)((j) - 1) * (m + 1) + (i) - 1]
    memcpy(p1, sol, sizeof(double) * dim);
    for (int i = 0; i < o->nc; ++i)
        for (int ib = 0; ib < 5; ++ib) sol[(size_t)5 * i + ib] = p1[(size_t)5 * i + ib] / diag[(size_t)25 * i + ib * 6];
    double ref = 0.0;
    for (size_t i = 0; i < dim; ++i) ref += fabs(sol[i]);
    const double normp1 = sqrt(vdot(p1, p1, dim));
    double errgmres = DBL_MAX;
    int nit = 0;
    while (errgmres >= ref * maxres && nit <= max_it * 2) {
        bdlu_yeqax(o, diag, upper, lower, sol, r1);
        for (size_t i = 0; i < dim; ++i) r1[i] = p1[i] - r1[i];
        const double beta = sqrt(vdot(r1, r1, dim));
        double fact = 1.0 / beta;
        for (size_t i = 0; i < dim; ++i) v1[i] = fact * r1[i];
        for (int i = 0; i <= m; ++i) ss[i] = 0.0;
        ss[0] = beta;
        for (size_t i = 0; i < (size_t)(m + 1) * m; ++i) hh[i] = 0.0;
        int jk;
        for (jk = 1; jk <= m; ++jk) {
            bdlu_yeqax(o, diag, upper, lower, v1 + (size_t)(jk - 1) * dim, w1);
            for (int ik = 1; ik <= jk; ++ik) {
                const double* vi = v1 + (size_t)(ik - 1) * dim;
                HH(ik, jk) = vdot(w1, vi, dim);
                for (size_t i = 0; i < dim; ++i) w1[i] = w1[i] - HH(ik, jk) * vi[i];
            }
            HH(jk + 1, jk) = sqrt(vdot(w1, w1, dim));
            for (int ik = 1; ik <= jk - 1; ++ik) {
                double tmp = HH(ik, jk);
                HH(ik, jk) = qc[ik - 1] * tmp + qs[ik - 1] * HH(ik + 1, jk);
                HH(ik + 1, jk) = -qs[ik - 1] * tmp + qc[ik - 1] * HH(ik + 1, jk);
            }
            if (HH(jk + 1, jk) > DBL_EPSILON * 10 * normp1) {
                if (jk < m) {
                    fact = 1.0 / HH(jk + 1, jk);
                    double* vn = v1 + (size_t)jk * dim;
                    for (size_t i = 0; i < dim; ++i) vn[i] = fact * w1[i];
                }
                if (HH(jk, jk) == 0) break;
                double tmp = sqrt(HH(jk, jk) * HH(jk, jk) + HH(jk + 1, jk) * HH(jk + 1, jk));
                qc[jk - 1] = HH(jk, jk) / tmp;
                qs[jk - 1] = HH(jk + 1, jk) / tmp;
                HH(jk, jk) = tmp;
                HH(jk + 1, jk) = 0.0;
                ss[jk] = -qs[jk - 1] * ss[jk - 1];
                ss[jk - 1] = qc[jk - 1] * ss[jk - 1];
                if (fabs(ss[jk]) < maxres * fabs(ss[0])) break;
            } else {
                break;
            }
        }
        const int nk = jk < m ? jk : m;
        for (int j = nk; j >= 1; --j) {
            double tmp = ss[j - 1];
            for (int ik = j + 1; ik <= nk; ++ik) tmp = tmp - HH(j, ik) * yy[ik - 1];
            yy[j - 1] = tmp / HH(j, j);
            const double* vj = v1 + (size_t)(j - 1) * dim;
            for (size_t i = 0; i < dim; ++i) sol[i] = sol[i] + yy[j - 1] * vj[i];
        }
        bdlu_yeqax(o, diag, upper, lower, sol, r1);
        for (size_t i = 0; i < dim; ++i) r1[i] = p1[i] - r1[i];
        ss[nk] = sqrt(vdot(r1, r1, dim));
        errgmres = ss[nk] / normp1;
        nit = nit + 1;
    }
#undef HH
    free(p1); free(r1); free(v1); free(w1); free(hh); free(ss); free(qc); free(qs); free(yy);
    return nit <= max_it ? nit - 1 : -1;
}

/* one implicit time step of the reference (calc_rhs with Jacobians, tstep_implicit.f90 ->
   implicit_solve.f90:39-84 with the GMRES method, update_field.f90:38): the residual of every
   cell including the ghost rows (their boundary-flux sums, flux_to_res.f90:90-113), the assembled
   matrix, bdlu_gmres, U += solution on interior cells.  Returns the GMRES info. */
int or_implicit_bdlu_step(or_ctx* o, double dt, int nkrylov, double maxres, int max_it)
{
    calc_rhs_internal(o, dt);
    double* diag = (double*)malloc(sizeof(double) * 25 * (size_t)o->nc);
    double* upper = (double*)malloc(sizeof(double) * 25 * (size_t)o->nf);
    double* lower = (double*)malloc(sizeof(double) * 25 * (size_t)o->nf);
    double* sol = (double*)malloc(sizeof(double) * 5 * (size_t)o->nc);
    or_bdlu_assemble(o, dt, diag, upper, lower);
    for (int i = 0; i < o->nc; ++i)
        for (int v = 0; v < NV; ++v) sol[(size_t)5 * i + v] = RES(o, v, i);
    int info = or_bdlu_gmres(o, diag, upper, lower, sol, nkrylov, maxres, max_it);
    for (int i = 0; i < o->ni; ++i)
        for (int v = 0; v < NV; ++v) CONS(o, v, i) = CONS(o, v, i) + sol[(size_t)5 * i + v];
    free(diag); free(upper); free(lower); free(sol);
    return info;
}

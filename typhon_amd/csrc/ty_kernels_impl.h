// HIP kernels of the residual path for gfx950 (MI355X).
//
// Layout: one work-item per interior cell (Hilbert or RCM order).  Each cell walks
// its face slots (ELL, [slot][cell] so a wavefront's slot loads coalesce) and
// evaluates every face flux in the face's canonical orientation (owner = the
// reference's ic1 side, geometry in the owner's frame), so the two cells of a
// face compute bit-identical fluxes.  The residual is a deterministic per-cell
// sum in slot order: no atomics, no face-flux array.  Geometry is streamed per
// slot; the only gathers are whole records (a neighbour's state, the far side's
// face state), each one or two cache lines.  Fluxes and reconstruction follow the
// reference routines cited per function.
#pragma once
#include "ty_internal.h"

namespace ty {

// AMD flang's MIN/MAX (a<b?a:b, a>b?a:b): second operand unless the first wins.
template <typename R>
__device__ __forceinline__ R fmin_(R a, R b) { return a < b ? a : b; }
template <typename R>
__device__ __forceinline__ R fmax_(R a, R b) { return a > b ? a : b; }

__device__ __forceinline__ double dpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float dpow(float a, float b) { return powf(a, b); }


// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD
// owns one contiguous range of RCM-ordered cells and its L2 sees the
// neighbours of its own cells.
__device__ __forceinline__ int xcd_block()
{
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// store_prim (ty_internal.h) with the state's dimension known at compile time
template <int DIM, typename R>
__device__ __forceinline__ void store_prim_n(const double* u, double g1, R* out)
{
    constexpr int NV = DIM + 2;
    const double rho = u[0];
    const double ir = 1.0 / rho;
    double vel[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) vel[d] = u[2 + d] * ir;
    double sq = vel[0] * vel[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) sq += vel[d] * vel[d];
    R q[NV];
    q[0] = (R)rho;
    q[1] = (R)(g1 * (u[1] - .5 * rho * sq));
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[2 + d] = (R)vel[d];
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = q[v];
}

// the cell's primitive record (rho, p, u..) and ir = 1/rho
template <int DIM, typename R>
__device__ __forceinline__ void load_prim(const R* __restrict__ us, int c, R* q, R& ir)
{
    constexpr int NV = DIM + 2;
    const R* u = us + (size_t)c * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = u[v];
    ir = arcp(q[0]);
}

template <int DIM, typename R>
__device__ __forceinline__ void load_prim(const R* __restrict__ us, int c, R* q)
{
    constexpr int NV = DIM + 2;
    const R* u = us + (size_t)c * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = u[v];
}

template <int DIM, typename R>
__device__ __forceinline__ R dotd(const R* a, const R* b)
{
    R s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) s += a[d] * b[d];
    return s;
}

// dynamic viscosity (calc_viscosity, EQNS.f90:225-265) in the reference's form
template <typename R>
__device__ __forceinline__ R visc_mu(const Physics& ph, R rho, R T)
{
    if (ph.visc_law == 11)
        return R(ph.visc_dyn) * dpow(T / R(ph.tref), R(1.5)) * (R(ph.tref) + R(ph.tsuth)) / (T + R(ph.tsuth));
    if (ph.visc_law == 14) return R(ph.visc_kin) * rho;
    return R(ph.visc_dyn);
}

// Boundary states from the interior cell's (cell_state) or the left face state
// (face_state, face = true); n is the outward unit normal, rfc = x_f - x_c of the
// interior cell (read by the flux wall in cell_state mode only):
// symmetry calcboco_ust_sym.f90:46-79, supersonic inlet setboco_ns_inlet_sup.f90
// (state precomputed on the host from pi_ti_mach_dir2nspri, EQNS.f90:139-160),
// subsonic inlet setboco_ns_inlet_sub.f90:75-134 (uniform PI/TI/DIRECTION,
// pi_ti_ps_dir2nspri EQNS.f90:166-191), subsonic outlet setboco_ns_outlet_sub.f90:57-95
// (EQNS.f90:166-219), supersonic outlet setboco_ns_outlet_sup.f90 (copy), isothermal wall
// setboco_ns_isoth.f90:66-84, flux and adiabatic walls setboco_ns_flux.f90:79-141.
template <int DIM, typename R>
__device__ __forceinline__ void bc_state(const BocoDev& b, const R* n, const R* qi, R* qo, const Physics& ph,
                                         bool face = false, const R* rfc = nullptr)
{
    constexpr int NV = DIM + 2;
    switch (b.kind) {
    case BK_WALL_ISOTH: {
        // T_g = 2 T_w - p/(rho r), p_g = p, rho_g = p_g/(r T_g), u_g = 2 u_w - u
        const R r = R(ph.rgas);
        const R temp = R(2) * R(b.param[0]) - qi[1] / (qi[0] * r);
        qo[1] = qi[1];
        qo[0] = qo[1] / (r * temp);
#pragma unroll
        for (int d = 0; d < DIM; ++d) qo[2 + d] = R(2) * R(b.param[4 + d]) - qi[2 + d];
        break;
    }
    case BK_WALL_FLUX: {
        const R r = R(ph.rgas);
        if (face) {
            // face_state: the reference's weak condition rho_R = p_L (setboco_ns_flux.f90:136)
            qo[0] = qi[1];
            qo[1] = qi[1];
        } else {
            // cell_state: T_g = T - flux d / conduct, d = |x_f - x_c|^2 / |(x_f - x_c).n|
            const R dd = dotd<DIM>(rfc, rfc) / fabs(dotd<DIM>(rfc, n));
            const R TH = qi[1] / (r * qi[0]);
            const R conduct = visc_mu(ph, qi[0], TH) * R(ph.cp_pr);
            const R temp = qi[1] / (qi[0] * r) - R(b.param[0]) * dd / conduct;
            qo[1] = qi[1];
            qo[0] = qo[1] / (r * temp);
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) qo[2 + d] = R(2) * R(b.param[4 + d]) - qi[2 + d];
        break;
    }
    case BK_INLET_SUB: {
        const R g = R(ph.gamma), g1 = g - R(1);
        const R ps = qi[1];
        R fm = dpow(R(b.param[0]) / ps, g1 / g);
        if (fm < R(1)) fm = R(1);
        const R rho = ps / (R(ph.rgas) * R(b.param[1])) * fm;
        const R vm = sqrt((fm - R(1)) * R(2) / g1) * sqrt(g * ps / rho);
        qo[0] = rho;
        qo[1] = ps;
#pragma unroll
        for (int d = 0; d < DIM; ++d) qo[2 + d] = vm * R(b.param[3 + d]);
        break;
    }
    case BK_SYM: {
        const R vn = dotd<DIM>(qi + 2, n);
        const R k = R(2) * vn;
        qo[0] = qi[0];
        qo[1] = qi[1];
#pragma unroll
        for (int d = 0; d < DIM; ++d) qo[2 + d] = qi[2 + d] - k * n[d];
        break;
    }
    case BK_INLET_SUP:
#pragma unroll
        for (int v = 0; v < NV; ++v) qo[v] = R(b.param[v]);
        break;
    case BK_OUTLET_SUB: {
        const R g = R(ph.gamma), g1 = g - R(1), r = R(ph.rgas);
        R fmv = dotd<DIM>(qi + 2, qi + 2);
        const R sq = sqrt(fmv);
        R dir[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) dir[d] = qi[2 + d] / sq;
        fmv = fmv / (g * qi[1]) * qi[0];
        fmv = R(1) + R(.5) * g1 * fmv;
        const R pi = qi[1] * dpow(fmv, g / g1);
        const R ti = qi[1] / r / qi[0] * fmv;
        const R ps = R(b.param[0]);
        R fm = dpow(pi / ps, g1 / g);
        if (fm < R(1)) fm = R(1);
        const R rho = ps / (r * ti) * fm;
        const R vm = sqrt((fm - R(1)) * R(2) / g1) * sqrt(g * ps / rho);
        qo[0] = rho;
        qo[1] = ps;
#pragma unroll
        for (int d = 0; d < DIM; ++d) qo[2 + d] = vm * dir[d];
        break;
    }
    default:
#pragma unroll
        for (int v = 0; v < NV; ++v) qo[v] = qi[v];
    }
}

// ghost-cell gradients for the viscous flux (cell_grad mode): symmetry mirrors
// (calcboco_ust_sym.f90:65-77 with calc_symn, VEC3D.f90:679-694); isothermal wall: grad p
// negated and mirrored, the rest copied (setboco_ns_isoth.f90:92-105); flux/adiabatic walls and
// the subsonic inlet: zero, as their loops (setboco_ns_flux.f90:156, setboco_ns_inlet_sub.f90:146,
// bound iend(ifb) with ifb unset) run no iteration in the reference; other inlets and outlets copy.
template <int DIM, typename R>
__device__ __forceinline__ void bc_grad(const BocoDev& b, const R* n, const R* gi, R* go)
{
    constexpr int NV = DIM + 2;
    if (b.kind != BK_SYM) {
#pragma unroll
        for (int k = 0; k < NV * DIM; ++k) go[k] = gi[k];
        if (b.kind == BK_WALL_ISOTH) {
            R gp[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) gp[d] = -gi[DIM + d];
            const R k2 = R(2) * dotd<DIM>(gp, n);
#pragma unroll
            for (int d = 0; d < DIM; ++d) go[DIM + d] = gp[d] - k2 * n[d];
        } else if (b.kind == BK_WALL_FLUX || b.kind == BK_INLET_SUB) {
#pragma unroll
            for (int k = 0; k < NV * DIM; ++k) go[k] = R(0);
        }
        return;
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const R k2 = R(2) * dotd<DIM>(gi + v * DIM, n);
#pragma unroll
        for (int d = 0; d < DIM; ++d) go[v * DIM + d] = gi[v * DIM + d] - k2 * n[d];
    }
    R S[DIM][DIM], M[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) S[i][j] = (i == j ? R(1) : R(0)) - R(2) * n[i] * n[j];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            R s = R(0);
#pragma unroll
            for (int l = 0; l < DIM; ++l) s += gi[(2 + i) * DIM + l] * S[l][j];
            M[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            R s = R(0);
#pragma unroll
            for (int l = 0; l < DIM; ++l) s += S[i][l] * M[l][j];
            go[(2 + i) * DIM + j] = s;
        }
}

// ---------------------------------------------------------------------------
// numerical fluxes per unit area, primitive inputs (rho, p, u..)

// Roe averages (calc_roe_states.f90:35-53) from the primitive states and their
// reciprocal densities: density, velocity, total enthalpy h and the squared sound
// speed a2 = (gamma - 1)(h - |v|^2 / 2) (the reference's gamma p / rho of its averaged
// pressure p = (h - |v|^2/2) rho (gamma - 1) / gamma, without the two divisions).
template <int DIM, typename R>
__device__ __forceinline__ void roe_avg(R g, R gig1, const R* qL, const R* qR, R irL, R irR, R& rho, R* vel, R& h,
                                        R& a2)
{
    const R t1 = asqrt(qR[0] * irL);
    rho = qL[0] * t1;
    const R kl = arcp(R(1) + t1);
    const R kr = t1 * kl;
#pragma unroll
    for (int d = 0; d < DIM; ++d) vel[d] = kl * qL[2 + d] + kr * qR[2 + d];
    const R hl = gig1 * qL[1] * irL + R(.5) * dotd<DIM>(qL + 2, qL + 2);
    const R hr = gig1 * qR[1] * irR + R(.5) * dotd<DIM>(qR + 2, qR + 2);
    h = kl * hl + kr * hr;
    a2 = (g - R(1)) * (h - R(.5) * dotd<DIM>(vel, vel));
}

// Wave speeds of the HLL family (calc_flux_hlle.f90:64-123, calc_flux_hllc.f90:64-125): Einfeldt's
// bounds from the sides' normal velocities and sound speeds and the Roe average (vm, am), with
// beta = 1 or sqrt((g-1)/2g) (the B variants), HLLE's bounded by zero; the kinetic variants (23, 24,
// 27, 28) take the mean speeds of each side's half Maxwellian, c h(M), c = a / sqrt(g/2),
// M = vn / c -+ sqrt(g/2) beta, in place of vn -+ beta a.
template <typename R>
__device__ __forceinline__ void hll_speeds(int sch, R g, R vnl, R vnr, R al, R ar, R vm, R am, R& sl, R& sr)
{
    const bool kin = sch == 23 || sch == 24 || sch == 27 || sch == 28;
    const R beta = (sch == 22 || sch == 24 || sch == 26 || sch == 28) ? sqrt((g - R(1)) / (R(2) * g)) : R(1);
    if (kin) {
        const R f1 = sqrt(R(.5) * g), f2 = R(.5) / sqrt(R(3.141592653589793238462643));
        const R cl = al / f1, cr = ar / f1;
        const R ml = vnl / cl - f1 * beta, mr = vnr / cr + f1 * beta;
        const R kinwl = R(.5) * (R(1) - erf(ml)), kinwr = R(.5) * (R(1) + erf(mr));
        const R kindl = -(f2 * exp(-ml * ml)), kindr = f2 * exp(-mr * mr);
        sl = fmin_(cl * (ml * kinwl + kindl), vm - am);
        sr = fmax_(cr * (mr * kinwr + kindr), vm + am);
    } else if (sch == 21 || sch == 22) {
        sl = fmin_(fmin_(R(0), vnl - beta * al), vm - am);
        sr = fmax_(fmax_(R(0), vnr + beta * ar), vm + am);
    } else {
        sl = fmin_(vnl - beta * al, vm - am);
        sr = fmax_(vnr + beta * ar, vm + am);
    }
}

// ig1 = 1/(gamma - 1), gig1 = gamma/(gamma - 1), irL/irR the reciprocal densities
template <int DIM, typename R>
__device__ __forceinline__ void num_flux_pre(int sch, R g, R ig1, R gig1, const R* qL, const R* qR, R irL, R irR,
                                             const R* n, R* F)
{
    const R vnl = dotd<DIM>(qL + 2, n);
    const R vnr = dotd<DIM>(qR + 2, n);
    const R rel = ig1 * qL[1] + R(.5) * qL[0] * dotd<DIM>(qL + 2, qL + 2);
    const R rer = ig1 * qR[1] + R(.5) * qR[0] * dotd<DIM>(qR + 2, qR + 2);
    const R al = asqrt(g * qL[1] * irL), ar = asqrt(g * qR[1] * irR);
    if (sch == 5) {  // Rusanov: calc_flux_rusanov.f90:64-83
        const R ray = fmax_(fabs(vnl) + al, fabs(vnr) + ar);
        F[0] = R(.5) * ((vnl + ray) * qL[0] + (vnr - ray) * qR[0]);
        F[1] = R(.5) * ((vnl + ray) * rel + (vnr - ray) * rer + vnl * qL[1] + vnr * qR[1]);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            F[2 + d] = R(.5) * (((vnl + ray) * qL[0]) * qL[2 + d] + ((vnr - ray) * qR[0]) * qR[2 + d]
                                + (qL[1] + qR[1]) * n[d]);
        return;
    }
    R rrho, rv[DIM], H, a2;
    roe_avg<DIM>(g, gig1, qL, qR, irL, irR, rrho, rv, H, a2);
    const R vm = dotd<DIM>(rv, n);
    const R am = asqrt(a2);
    if (sch >= 21 && sch <= 24) {  // HLLE, HLLEB, HLLEK, HLLEKB: calc_flux_hlle.f90:109-143
        R sl, sr;
        hll_speeds(sch, g, vnl, vnr, al, ar, vm, am, sl, sr);
        const R iks = arcp(sr - sl);
        const R kl = sr * iks, kr = -sl * iks, ku = sl * sr * iks;
        F[0] = (kl * vnl - ku) * qL[0] + (kr * vnr + ku) * qR[0];
        F[1] = (kl * vnl - ku) * rel + (kr * vnr + ku) * rer + (kl * vnl * qL[1] + kr * vnr * qR[1]);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            F[2 + d] = ((kl * vnl - ku) * qL[0]) * qL[2 + d] + ((kr * vnr + ku) * qR[0]) * qR[2 + d]
                       + (kl * qL[1] + kr * qR[1]) * n[d];
        return;
    }
    if (sch == 10) {  // Roe with Harten entropy fix (the reference parses ROE but stops on it)
        const R a = am;
        const R drho = qR[0] - qL[0], dp = qR[1] - qL[1], dvn = vnr - vnl;
        R l1 = fabs(vm - a), l2 = fabs(vm), l3 = fabs(vm + a);
        const R del = R(0.1) * a;
        if (l1 < del) l1 = R(.5) * (l1 * l1 / del + del);
        if (l3 < del) l3 = R(.5) * (l3 * l3 / del + del);
        const R w1 = l1 * (dp - rrho * a * dvn) / (R(2) * a * a);
        const R w2 = l2 * (drho - dp / (a * a));
        const R w3 = l3 * (dp + rrho * a * dvn) / (R(2) * a * a);
        const R w4 = l2 * rrho;
        R dvt[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) dvt[d] = (qR[2 + d] - qL[2 + d]) - dvn * n[d];
        F[0] = R(.5) * (qL[0] * vnl + qR[0] * vnr - (w1 + w2 + w3));
        F[1] = R(.5) * (vnl * (rel + qL[1]) + vnr * (rer + qR[1])
                        - (w1 * (H - vm * a) + w2 * R(.5) * dotd<DIM>(rv, rv) + w3 * (H + vm * a)
                           + w4 * dotd<DIM>(rv, dvt)));
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            F[2 + d] = R(.5) * (qL[0] * vnl * qL[2 + d] + qL[1] * n[d] + qR[0] * vnr * qR[2 + d] + qR[1] * n[d]
                                - (w1 * (rv[d] - a * n[d]) + w2 * rv[d] + w3 * (rv[d] + a * n[d]) + w4 * dvt[d]));
        return;
    }
    // HLLC (Batten), HLLCB, HLLCK, HLLCKB: calc_flux_hllc.f90:61-171
    R sl, sr;
    hll_speeds(sch, g, vnl, vnr, al, ar, vm, am, sl, sr);
    if (sl >= R(0)) {
        F[0] = vnl * qL[0];
        F[1] = vnl * (rel + qL[1]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[2 + d] = (vnl * qL[0]) * qL[2 + d] + qL[1] * n[d];
    } else if (sr <= R(0)) {
        F[0] = vnr * qR[0];
        F[1] = vnr * (rer + qR[1]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[2 + d] = (vnr * qR[0]) * qR[2 + d] + qR[1] * n[d];
    } else {
        const R rqL = qL[0] * (sl - vnl), rqR = qR[0] * (sr - vnr);
        const R sst = adiv(rqR * vnr - rqL * vnl - qR[1] + qL[1], rqR - rqL);
        // star state of the side of S* (calc_flux_hllc.f90:149-161)
        const bool left = sst >= R(0);
        const R* q = left ? qL : qR;
        const R rq = left ? rqL : rqR;
        const R vn = left ? vnl : vnr;
        const R s = left ? sl : sr;
        const R re = left ? rel : rer;
        const R iks = arcp(s - sst);
        const R rst = rq * iks;
        const R pst = q[1] - rq * (vn - sst);
        const R rest = iks * (re * (s - vn) - q[1] * vn + pst * sst);
        F[0] = sst * rst;
        F[1] = sst * (rest + pst);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const R rvst = iks * ((rq * q[2 + d]) + (pst - q[1]) * n[d]);
            F[2 + d] = (sst * rvst) + (pst * n[d]);
        }
    }
}

template <int DIM, typename R>
__device__ void num_flux(int sch, R g, const R* qL, const R* qR, const R* n, R* F)
{
    const R ig1 = R(1) / (g - R(1));
    const R gig1 = g / (g - R(1));
    num_flux_pre<DIM>(sch, g, ig1, gig1, qL, qR, R(1) / qL[0], R(1) / qR[0], n, F);
}

// ---------------------------------------------------------------------------
// reconstruction: limiters (SOURCE/MGRID/LIMITER.f90:72-339)

template <typename R>
__device__ __forceinline__ R limiter(char lim, R x1, R x2)
{
    R r = R(0);
    switch (lim) {
    case 'M': {
        const R s = R(.5) * (copysign(R(1), x1) + copysign(R(1), x2));
        r = s * fmin_(fabs(x1), fabs(x2));
        break;
    }
    case 'A': {
        const R s = x1 * x2;
        r = (s <= R(0)) ? R(0) : s * (x1 + x2) / (x1 * x1 + x2 * x2);
        break;
    }
    case 'V': {
        const R s = x1 * x2;
        r = (s <= R(0)) ? R(0) : R(2) * s / (x1 + x2);
        break;
    }
    case 'S':
        if (x1 > R(0) && x2 > R(0)) r = fmax_(fmin_(R(2) * x1, x2), fmin_(R(2) * x2, x1));
        if (x1 < R(0) && x2 < R(0)) r = fmin_(fmax_(R(2) * x1, x2), fmax_(R(2) * x2, x1));
        break;
    case 'K':
        if (x1 > R(0) && x2 > R(0)) { r = (R(2) * x1 + x2) / R(3); r = fmin_(fmin_(R(2) * x1, R(2) * x2), r); }
        if (x1 < R(0) && x2 < R(0)) { r = (R(2) * x1 + x2) / R(3); r = fmax_(fmax_(R(2) * x1, R(2) * x2), r); }
        break;
    case 'L': {
        const R eta = (x1 * x1 + x2 * x2) / R(0.01) / R(0.01);
        if (eta < R(1)) r = (R(2) * x2 + x1) / R(3);
        else if (x1 > R(0) && x2 > R(0)) { r = (R(2) * x2 + x1) / R(3); r = fmin_(fmin_(R(1.6) * x2, R(2) * x1), r); }
        else if (x1 < R(0) && x2 < R(0)) { r = (R(2) * x2 + x1) / R(3); r = fmax_(fmax_(R(1.6) * x2, R(2) * x1), r); }
        else if (x1 < R(0) && x2 > R(0)) { r = (R(2) * x2 + x1) / R(3); r = fmax_(R(0), fmin_(R(-0.5) * x1, r)); }
        else if (x1 > R(0) && x2 < R(0)) { r = (R(2) * x2 + x1) / R(3); r = fmin_(R(0), fmax_(R(-0.5) * x1, r)); }
        break;
    }
    default: break;
    }
    return r;
}

// One side's face value for the MUSCL family: hres_ns_muscl.f90:51-143 with
// tvdgradstr_*.f90:35-58, hres_ns_musclfast.f90:51-144 with tvdgradfst_*.f90:35-60,
// hres_ns_muscluns with tvdgraduns_*.f90:35-57.  g is this side's gradient of q,
// u the (possibly unnormalised) L->R vector, lr the L->R slope, cf the cell->face
// vector (MUSCL, MUSCL-UNS) and kcf its signed length (MUSCL-FAST).
template <int DIM, typename R>
__device__ __forceinline__ R muscl_side(char method, char lim, R q, const R* g, const R* u, R lr, const R* cf, R kcf)
{
    if (method == 'F') {
        R s;
        if (lim == 'X') s = lr;
        else {
            const R sp = dotd<DIM>(g, u);
            s = (lim == 'K') ? limiter(lim, lr, R(2) * sp - lr) : limiter(lim, R(2) * sp - lr, lr);
        }
        return q + s * kcf;
    }
    const R s = dotd<DIM>(g, u);
    R ds;
    if (lim == 'X') ds = lr - s;
    else if (method == 'M') ds = ((lim == 'K') ? limiter(lim, lr, R(2) * s - lr) : limiter(lim, R(2) * s - lr, lr)) - s;
    else ds = ((lim == 'K') ? limiter(lim, lr, s) : limiter(lim, s, lr)) - s;
    R gp[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) gp[d] = g[d] + ds * u[d];
    return q + dotd<DIM>(gp, cf);
}

// Venkatakrishnan smooth limiter function (J. Comput. Phys. 118 (1995) 120-130)
template <typename R>
__device__ __forceinline__ R venkat_phi(R dmax, R dmin, R d2, R eps2)
{
    if (d2 == R(0)) return R(1);
    const R dm = (d2 > R(0)) ? dmax : dmin;
    const R num = (dm * dm + eps2) + R(2) * d2 * dm;
    const R den = dm * dm + R(2) * d2 * d2 + dm * d2 + eps2;
    if (den <= R(0)) return R(1);
    return fmin_(R(1), fmax_(R(0), num / den));
}

// ---------------------------------------------------------------------------
// slot decoding and per-slot geometry

struct SlotInfo {
    int nb, f;
    bool per, other;
};

__device__ __forceinline__ SlotInfo decode(int w, int nb)
{
    SlotInfo s;
    s.nb = nb;
    s.f = (int)((unsigned)w & FACE_MASK);
    s.per = ((unsigned)w & PERIODIC_BIT) != 0u;
    s.other = w < 0;
    return s;
}

// one 3-vector of slot k of cell c from a [slot][3][cell] array
template <int DIM, typename R>
__device__ __forceinline__ void slot_vec(const R* __restrict__ p, int k, int n, int c, R* x)
{
#pragma unroll
    for (int e = 0; e < DIM; ++e) x[e] = p[((size_t)k * 3 + e) * n + c];
}

// slot_vec of a stream read once per launch, non-temporal (it need not displace gathered data)
template <int DIM, typename R>
__device__ __forceinline__ void slot_vec_nt(const R* __restrict__ p, int k, int n, int c, R* x)
{
#pragma unroll
    for (int e = 0; e < DIM; ++e) x[e] = __builtin_nontemporal_load(p + ((size_t)k * 3 + e) * n + c);
}

template <int DIM, typename R>
__device__ __forceinline__ void face_n(const Geo<R>& g, int f, R* n)
{
    n[0] = g.fnx[f];
    n[1] = g.fny[f];
    if (DIM == 3) n[DIM - 1] = g.fnz[f];
}

template <int K, typename R>
__device__ __forceinline__ void load_rec(const R* __restrict__ p, int i, R* out)
{
    const R* r = p + (size_t)i * K;
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = r[k];
}

template <int DIM, typename R>
__device__ __forceinline__ void load_grad(const R* __restrict__ grad, int c, R* gv)
{
    constexpr int K = (DIM + 2) * DIM;
    const R* g = grad + (size_t)c * K;
#pragma unroll
    for (int k = 0; k < K; ++k) gv[k] = g[k];
}

// Slot link word (per slot, per cell): bit 31 boundary face, bit 30 this cell is the far
// side of an interior face, bits 27-29 the slot where the far cell keeps the same face,
// bits 0-26 the neighbour cell or the boundary face index; SLOT_UNUSED for an empty slot.
constexpr int LINK_INDEX = (1 << 27) - 1;
__device__ __forceinline__ bool link_boundary(int L) { return L < 0; }
__device__ __forceinline__ bool link_other(int L) { return (L >> 30) & 1; }
__device__ __forceinline__ int link_index(int L) { return L & LINK_INDEX; }
__device__ __forceinline__ int link_mirror(int L) { return (L >> 27) & (MAX_SLOTS - 1); }

// unit normal and area of slot k's face from its canonical area vector S = A n
template <int DIM, typename R>
__device__ __forceinline__ R slot_normal(const Geo<R>& g, int k, int n, int c, R* nf)
{
    R S[DIM];
    slot_vec<DIM>(g.ss, k, n, c, S);
    const R A = asqrt(dotd<DIM>(S, S));
    const R ia = arcp(A);
#pragma unroll
    for (int e = 0; e < DIM; ++e) nf[e] = S[e] * ia;
    return A;
}

// The state across slot k: the neighbour's primitive state, or the ghost state of the
// cell_state boundary map (calcboco_ust.f90:44-105) built from this cell's.
template <int DIM, typename R>
__device__ __forceinline__ void ghost_state(const KArgs<R>& a, int k, int c, int L, const R* qc, R* qo)
{
    R nf[DIM];
    slot_normal<DIM>(a.g, k, a.n, c, nf);
    const BocoDev& b = a.boco[a.g.bboco[link_index(L)]];
    if (b.kind == BK_WALL_FLUX) {
        R rk[DIM];
        slot_vec<DIM>(a.g.sr, k, a.n, c, rk);
        bc_state<DIM>(b, nf, qc, qo, a.ph, false, rk);
    } else {
        bc_state<DIM>(b, nf, qc, qo, a.ph);
    }
}

// BND false: the caller knows the slot is not a boundary slot (no boundary code is generated)
template <int DIM, bool BND = true, typename R>
__device__ __forceinline__ void far_state(const KArgs<R>& a, int k, int c, int L, const R* qc, R* qo, R& ir)
{
    if (!BND || !link_boundary(L)) {
        load_prim<DIM>(a.us, link_index(L), qo, ir);
    } else {
        ghost_state<DIM>(a, k, c, L, qc, qo);
        ir = arcp(qo[0]);
    }
}

template <int DIM, bool BND = true, typename R>
__device__ __forceinline__ void far_state(const KArgs<R>& a, int k, int c, int L, const R* qc, R* qo)
{
    if (!BND || !link_boundary(L)) {
        load_prim<DIM>(a.us, link_index(L), qo);
    } else {
        ghost_state<DIM>(a, k, c, L, qc, qo);
    }
}

// One side's MUSCL-family face value from the canonical face data (hres_ns_muscl.f90:51-143
// and its variants): qL/qR are the canonical cell states, d = x_R - x_L, r = x_f - x_c in this
// cell's frame, g this cell's gradient; this cell is the left side when !other.
template <int DIM, typename R>
__device__ __forceinline__ void muscl_own(const Physics& ph, const R* qL, const R* qR, const R* d, const R* r,
                                          const R* g, bool other, R* qf)
{
    constexpr int NV = DIM + 2;
    R u[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) u[e] = d[e];
    R dlr, kcf = R(0);
    if (ph.method == 'F') {
        // x_f - x_L and x_f - x_R from this cell's r and the canonical d
        R LF[DIM], RF[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            LF[e] = other ? r[e] + d[e] : r[e];
            RF[e] = other ? r[e] : r[e] - d[e];
        }
        const R kl = sqrt(dotd<DIM>(LF, LF));
        const R kr = -sqrt(dotd<DIM>(RF, RF));
        dlr = kl - kr;
        kcf = other ? kr : kl;
    } else {
        dlr = sqrt(dotd<DIM>(u, u));
    }
#pragma unroll
    for (int e = 0; e < DIM; ++e) u[e] = u[e] / dlr;
    const R* q = other ? qR : qL;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const R lr = (qR[v] - qL[v]) / dlr;
        qf[v] = muscl_side<DIM>(ph.method, ph.limiter, q[v], g + v * DIM, u, lr, r, kcf);
    }
}

// ---------------------------------------------------------------------------
// K1: LSQ gradients (calc_gradient.f90:75-192; the normal-equation solve of
// precalc_grad_lsq.f90:60-155 + cholesky_solve MATRIX.f90:79-111 is folded into
// per-slot weights w_k = A^-1 d_k computed once on the host) and this cell's
// own-side face states for every face slot: MUSCL / MUSCL-FAST / MUSCL-UNS with
// the reference limiters, Barth post-limitation (postlimit_barth.f90:50-223,
// ratios of rho, p and the velocity vector), or the Venkatakrishnan cell limiter.
// Neighbour set = face neighbours incl. ghost cells from the cell_state boundary map.

// side_temp (below) ahead of its definition
template <int DIM, typename R>
__device__ __forceinline__ void side_temp(const Physics& ph, const R* q, R ir, const R* g, R& T, R* gT);

// The viscous flux reads a side's gradients only as the velocity gradient and the temperature
// gradient (calc_flux_viscous_cell.f90:96-109): K1 stores those per cell (dim^2 + dim values
// instead of nv dim) and, for the ghost side of each boundary face, the gradients of the
// cell_grad boundary map (bc_grad) with the ghost's temperature gradient.
// BND false: the cell has no boundary slot (no ghost records)
template <int DIM, typename R, bool BND = true>
__device__ __forceinline__ void write_visc_records(const KArgs<R>& a, int c, const R* qc, const R* gr, R* dst,
                                                   bool ghosts)
{
    constexpr int VR = DIM * DIM + DIM;
    R* go = dst + (size_t)c * VR;
#pragma unroll
    for (int k = 0; k < DIM * DIM; ++k) go[k] = gr[2 * DIM + k];
    R T, gT[DIM];
    side_temp<DIM>(a.ph, qc, arcp(qc[0]), gr, T, gT);
#pragma unroll
    for (int e = 0; e < DIM; ++e) go[DIM * DIM + e] = gT[e];
    if (!BND || !ghosts) return;
    for (int k = 0; k < a.ns; ++k) {
        const int L = a.link[(size_t)k * a.n + c];
        if (L == SLOT_UNUSED || !link_boundary(L)) continue;
        R nf[DIM], qo[DIM + 2], gb[(DIM + 2) * DIM];
        slot_normal<DIM>(a.g, k, a.n, c, nf);
        ghost_state<DIM>(a, k, c, L, qc, qo);
        bc_grad<DIM>(a.boco[a.g.bboco[link_index(L)]], nf, gr, gb);
        R To, gTo[DIM];
        side_temp<DIM>(a.ph, qo, arcp(qo[0]), gb, To, gTo);
        R* ob = a.gbnd_out + (size_t)link_index(L) * VR;
#pragma unroll
        for (int q = 0; q < DIM * DIM; ++q) ob[q] = gb[2 * DIM + q];
#pragma unroll
        for (int e = 0; e < DIM; ++e) ob[DIM * DIM + e] = gTo[e];
    }
}

// reconstruction performed by K1 after the gradient
enum Rec : int { REC_NONE = 0, REC_VENKAT = 1, REC_MUSCL = 2, REC_MUSCL_BARTH = 3, REC_MUSCL_MONO = 4 };

// monotonic0/1/2 (LIMITER.f90:427-477) of the face values x1 (left side) and x2 (right side)
// between the cell values a (left) and b (right): MONOTONIC0 keeps each in [a, b], MONOTONIC1 also
// bounds each by the pair's average, MONOTONIC2 keeps each between its cell value and (a + b) / 2.
// mono_left returns the limited x1, mono_right the limited x2 (0 and 2 read only their own side).
template <typename R>
__device__ __forceinline__ R mono_left(int m, R x1, R x2, R a, R b)
{
    if (m == 10) return a < b ? fmax_(a, fmin_(x1, b)) : fmin_(a, fmax_(x1, b));
    if (m == 11) {
        const R midx = R(.5) * (x1 + x2);
        return a < b ? fmax_(a, fmin_(fmin_(x1, midx), b)) : fmin_(a, fmax_(fmax_(x1, midx), b));
    }
    const R mid = R(.5) * (a + b);
    return a < b ? fmax_(a, fmin_(x1, mid)) : fmin_(a, fmax_(x1, mid));
}
template <typename R>
__device__ __forceinline__ R mono_right(int m, R x1, R x2, R a, R b)
{
    if (m == 10) return a < b ? fmin_(b, fmax_(x2, a)) : fmax_(b, fmin_(x2, a));
    if (m == 11) {
        const R midx = R(.5) * (x1 + x2);
        return a < b ? fmin_(b, fmax_(fmax_(x2, midx), a)) : fmax_(b, fmin_(fmin_(x2, midx), a));
    }
    const R mid = R(.5) * (a + b);
    return a < b ? fmin_(b, fmax_(x2, mid)) : fmax_(b, fmin_(x2, mid));
}

// postlimit_monotonic (postlimit_monotonic.f90:40-110) on this cell's face state q of slot k in K1:
// MONOTONIC0/2 from the two cell values; MONOTONIC1 of a boundary face with the ghost side's own
// extrapolation (the hres of the ghost cell with the cell_grad boundary map of its gradient), its
// interior faces are left to k_mono1, which sees both sides' face states.
template <int DIM, bool BND = true, typename R>
__device__ __forceinline__ void mono_own(const KArgs<R>& a, int k, int c, int L, bool other, const R* qc, const R* qo,
                                         const R* dk, const R* rk, const R* gr, R* q)
{
    constexpr int NV = DIM + 2;
    const int m = a.ph.mono;
    if (m == 11) {
        if (!BND || !link_boundary(L)) return;
        R nf[DIM], gb[NV * DIM], rg[DIM], x2[NV];
        (void)slot_normal<DIM>(a.g, k, a.n, c, nf);
        bc_grad<DIM>(a.boco[a.g.bboco[link_index(L)]], nf, gr, gb);
#pragma unroll
        for (int e = 0; e < DIM; ++e) rg[e] = rk[e] - dk[e];
        muscl_own<DIM>(a.ph, qc, qo, dk, rg, gb, true, x2);
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = mono_left(11, q[v], x2[v], qc[v], qo[v]);
        return;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
        q[v] = other ? mono_right(m, R(0), q[v], qo[v], qc[v]) : mono_left(m, q[v], R(0), qc[v], qo[v]);
}

// DENSE: every cell uses all NSM == ns slots (closed tet/hex/quad meshes), so the slot
// loops carry no guards and every slot's loads can be issued together.
// GOUT: the cell gradients (or the viscous records) are written too; the Euler reconstructions
// launch without it, so they do not carry the viscous-record code's registers.
// Kernels that take the interior path (a copy of the cell's code without the boundary branches) for
// wavefronts without boundary slots: measured faster for the FP32 kernels and the plain-gradient
// and viscous FP64 kernels, slower for the FP64 Euler second-order pair.
template <typename R, int REC>
constexpr bool grad_interior_path() { return sizeof(R) == 4 || REC == REC_NONE; }
template <typename R, bool VISC>
constexpr bool flux_interior_path() { return sizeof(R) == 4 || VISC; }

// a streamed record read once (non-temporal: it need not displace gathered data in the caches)
__device__ __forceinline__ float4 nt_load4(const float4* p)
{
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 nt_load2(const double2* p)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_double2(v.x, v.y);
}
// record i of a packed slot array (Pack4): FP32 one float4, FP64 two double2 planes
__device__ __forceinline__ float4 ld_pack(const float4* p, size_t i, size_t) { return nt_load4(p + i); }
__device__ __forceinline__ double4 ld_pack(const double2* p, size_t i, size_t plane)
{
    const double2 a = nt_load2(p + i), b = nt_load2(p + plane + i);
    return make_double4(a.x, a.y, b.x, b.y);
}
// the link word stored in the fourth element of a packed record
__device__ __forceinline__ int link_of(float w) { return __float_as_int(w); }
__device__ __forceinline__ int link_of(double w) { return (int)__double_as_longlong(w); }

// Whether the wavefront of cell c holds a cell with a boundary slot: bits 2w and 2w + 1 (its two
// 32-cell groups) of the group bits, w = c / 64; uniform over the wavefront.
__device__ __forceinline__ bool wave_boundary(const unsigned* bits, int c)
{
    if (!bits) return true;
    const int w = __builtin_amdgcn_readfirstlane(c >> 6);
    return (bits[w >> 4] & (3u << ((2 * w) & 31))) != 0u;
}

template <int DIM, typename R, int NSM, int REC, bool DENSE, bool GOUT, bool BND>
__device__ __forceinline__ void grad_cell(const KArgs<R>& a, int c);

// Meshes whose cells use every slot run the wavefronts without boundary slots through a copy of
// the cell's code without the boundary branches, so that every slot's loads can be issued at once.
template <int DIM, typename R, int NSM, int REC, bool DENSE, bool GOUT>
__global__ __launch_bounds__(BLOCK) void k_gradient(KArgs<R> a)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n || (a.stop && *a.stop)) return;
    if (!DENSE || !grad_interior_path<R, REC>() || wave_boundary(a.bnd_bits, c))
        grad_cell<DIM, R, NSM, REC, DENSE, GOUT, true>(a, c);
    else grad_cell<DIM, R, NSM, REC, DENSE, GOUT, false>(a, c);
}

template <int DIM, typename R, int NSM, int REC, bool DENSE, bool GOUT, bool BND>
__device__ __forceinline__ void grad_cell(const KArgs<R>& a, int c)
{
    constexpr int NV = DIM + 2;
    const int n = a.n;
    R qc[NV];
    load_prim<DIM>(a.us, c, qc);
    R gr[NV * DIM];
    R qmax[NV], qmin[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        qmax[v] = qc[v];
        qmin[v] = qc[v];
#pragma unroll
        for (int e = 0; e < DIM; ++e) gr[v * DIM + e] = R(0);
    }
    // interior path: the link word and the LSQ weight in one packed record (pw)
    constexpr bool PW = DENSE && !BND;
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (!DENSE && k >= a.ns) break;
        decltype(ld_pack(a.g.pw, 0, 0)) pw;
        if (PW) pw = ld_pack(a.g.pw, (size_t)k * n + c, (size_t)a.ns * n);
        const int L = PW ? link_of(pw.w) : a.link[(size_t)k * n + c];
        if (!DENSE && L == SLOT_UNUSED) continue;
        R qo[NV], wk[DIM];
        far_state<DIM, BND>(a, k, c, L, qc, qo);
        if (PW) {
            const R w3[3] = {R(pw.x), R(pw.y), R(pw.z)};
#pragma unroll
            for (int e = 0; e < DIM; ++e) wk[e] = w3[e];
        } else if (DENSE) {
            slot_vec_nt<DIM>(a.g.sw, k, n, c, wk);
        } else {
            slot_vec<DIM>(a.g.sw, k, n, c, wk);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R dq = qo[v] - qc[v];
#pragma unroll
            for (int e = 0; e < DIM; ++e) gr[v * DIM + e] += wk[e] * dq;
            if (REC == REC_VENKAT) {
                qmax[v] = fmax_(qmax[v], qo[v]);
                qmin[v] = fmin_(qmin[v], qo[v]);
            }
        }
    }
    if (GOUT && a.grad_out) {
        if (a.vrec) write_visc_records<DIM, R, BND>(a, c, qc, gr, a.grad_out, true);
        else {
            R* go = a.grad_out + (size_t)c * NV * DIM;
#pragma unroll
            for (int k = 0; k < NV * DIM; ++k) go[k] = gr[k];
            if (a.vrec_out) write_visc_records<DIM, R, BND>(a, c, qc, gr, a.vrec_out, false);
        }
    }
    if (REC == REC_NONE) return;
    if (REC == REC_VENKAT) {
        // Venkatakrishnan: phi_v = min over faces, then q_f = q_c + phi grad(q).(x_f - x_c).
        // The face values phi_f = num/den lie in [0, 1] with den > 0 (venkat_phi), so the
        // minimum is tracked as a fraction by cross-multiplication and divided once.
        const R eps2 = a.g.eps2[c];
        R pn[NV], pd[NV];
        R d2k[NSM][NV];  // grad(q_v).(x_f - x_c) of each slot, kept for the face states
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            pn[v] = R(1);
            pd[v] = R(1);
        }
#pragma unroll
        for (int k = 0; k < NSM; ++k) {
            if (!DENSE && k >= a.ns) break;
            if (!DENSE && a.link[(size_t)k * n + c] == SLOT_UNUSED) continue;
            R rk[DIM];
            if (DENSE) slot_vec_nt<DIM>(a.g.sr, k, n, c, rk);
            else slot_vec<DIM>(a.g.sr, k, n, c, rk);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R d2 = dotd<DIM>(gr + v * DIM, rk);
                d2k[k][v] = d2;
                if (d2 == R(0)) continue;
                const R dm = (d2 > R(0)) ? qmax[v] - qc[v] : qmin[v] - qc[v];
                const R num = (dm * dm + eps2) + R(2) * d2 * dm;
                const R den = dm * dm + R(2) * d2 * d2 + dm * d2 + eps2;
                if (num * pd[v] < pn[v] * den) {
                    pn[v] = num;
                    pd[v] = den;
                }
            }
        }
        R phi[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) phi[v] = adiv(pn[v], pd[v]);
#pragma unroll
        for (int k = 0; k < NSM; ++k) {
            if (!DENSE && k >= a.ns) break;
            if (!DENSE && a.link[(size_t)k * n + c] == SLOT_UNUSED) continue;
            R* out = a.qf_out + ((size_t)k * n + c) * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) out[v] = qc[v] + phi[v] * d2k[k][v];
        }
        return;
    }
    // MUSCL family, optionally Barth-limited or monotonic post-limited
    constexpr bool barth = REC == REC_MUSCL_BARTH;
    const R klim = R(a.ph.klim);
    R ratio[3] = {R(1), R(1), R(1)};
    R qfk[barth ? NSM : 1][NV];
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (!DENSE && k >= a.ns) break;
        const int L = a.link[(size_t)k * n + c];
        if (!DENSE && L == SLOT_UNUSED) continue;
        const bool other = link_other(L);
        R qo[NV], dk[DIM], rk[DIM];
        far_state<DIM, BND>(a, k, c, L, qc, qo);
        slot_vec<DIM>(a.g.sd, k, n, c, dk);
        slot_vec<DIM>(a.g.sr, k, n, c, rk);
        R* q = qfk[barth ? k : 0];
        muscl_own<DIM>(a.ph, other ? qo : qc, other ? qc : qo, dk, rk, gr, other, q);
        if (barth) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const R l = (qo[v] - qc[v]) / (q[v] - qc[v]);
                const R lr = fmax_(klim * l, R(0));
                const int gi = v < 2 ? v : 2;
                ratio[gi] = fmin_(ratio[gi], lr);
            }
        } else {
            if (REC == REC_MUSCL_MONO) mono_own<DIM, BND>(a, k, c, L, other, qc, qo, dk, rk, gr, q);
            R* out = a.qf_out + ((size_t)k * n + c) * NV;
#pragma unroll
            for (int v = 0; v < NV; ++v) out[v] = q[v];
        }
    }
    if (!barth) return;
    // in 2D the reference also limits the (identically zero) z velocity: 0/0 -> max(NaN,0) = 0
    if (DIM == 2) ratio[2] = fmin_(ratio[2], R(0));
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (!DENSE && k >= a.ns) break;
        if (!DENSE && a.link[(size_t)k * n + c] == SLOT_UNUSED) continue;
        R* out = a.qf_out + ((size_t)k * n + c) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = qc[v] + ratio[v < 2 ? v : 2] * (qfk[barth ? k : 0][v] - qc[v]);
    }
}

// ---------------------------------------------------------------------------
// viscous flux of one face, calc_flux_viscous_cell.f90:75-158 (subtracted from F):
// qL/qR the canonical cell states, dHL the distance weight |x_f - x_L| / (|x_f - x_L| + |x_f - x_R|)
// (dHR = 1 - dHL), vLR = x_R - x_L, gL/gR the cell gradients, nf the canonical normal.
template <int DIM, typename R, typename W>
__device__ void viscous_face(const Physics& ph, const R* qL, const R* qR, W dHLw, const R* vLR,
                             const R* nf, const R* gL, const R* gR, R* flux)
{
    const R r = R(ph.rgas);
    const R cp = R(ph.gamma) * r / (R(ph.gamma) - R(1));
    const R dHL = R(dHLw), dHR = R(1.0 - dHLw);
    R velH[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) velH[e] = dHR * qL[2 + e] + dHL * qR[2 + e];
    const R TL = qL[1] / (qL[0] * r), TR = qR[1] / (qR[0] * r);
    const R TH = dHR * TL + dHL * TR;
    const R rhoH = dHR * qL[0] + dHL * qR[0];
    R gTL[DIM], gTR[DIM];
    const R kl1 = R(1) / (qL[0] * r), kl2 = qL[1] / (qL[0] * qL[0] * r);
    const R kr1 = R(1) / (qR[0] * r), kr2 = qR[1] / (qR[0] * qR[0] * r);
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        gTL[e] = kl1 * gL[DIM + e] - kl2 * gL[e];
        gTR[e] = kr1 * gR[DIM + e] - kr2 * gR[e];
    }
    R mu;
    if (ph.visc_law == 11) mu = R(ph.visc_dyn) * dpow(TH / R(ph.tref), R(1.5)) * (R(ph.tref) + R(ph.tsuth)) / (TH + R(ph.tsuth));
    else if (ph.visc_law == 14) mu = R(ph.visc_kin) * rhoH;
    else mu = R(ph.visc_dyn);
    R G[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) G[i][j] = dHL * gR[(2 + i) * DIM + j] + dHR * gL[(2 + i) * DIM + j];
    const R dLR2 = dotd<DIM>(vLR, vLR);
    R gavg[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) gavg[e] = dHL * gTR[e] + dHR * gTL[e];
    R gTH;
    if (ph.sch_dis == 10) {
        R kv[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) kv[e] = (R(1) / dLR2) * vLR[e];
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const R fa = qR[2 + i] - qL[2 + i] - dotd<DIM>(G[i], vLR);
#pragma unroll
            for (int j = 0; j < DIM; ++j) G[i][j] = G[i][j] + fa * kv[j];
        }
        const R pscal = dotd<DIM>(vLR, nf) / dLR2;
        const R fcomp = pscal * (TR - TL);
        R vi[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) vi[e] = nf[e] - pscal * vLR[e];
        gTH = fcomp + dotd<DIM>(gavg, vi);
    } else {
        gTH = dotd<DIM>(gavg, nf);
    }
    R tr = G[0][0];
#pragma unroll
    for (int i = 1; i < DIM; ++i) tr += G[i][i];
    const R addv = -(R(2) / R(3)) * tr;
    R sn[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            R sij = G[i][j] + G[j][i];
            if (i == j) sij += addv;
            s += (mu * sij) * nf[j];
        }
        sn[i] = s;
    }
    const R svn = dotd<DIM>(sn, velH);
#pragma unroll
    for (int e = 0; e < DIM; ++e) flux[2 + e] = flux[2 + e] - sn[e];
    flux[1] = flux[1] - svn;
    const R conduct = mu * cp / R(ph.prandtl);
    flux[1] = flux[1] - conduct * gTH;
}

// One side's temperature and temperature gradient for the viscous flux
// (calc_flux_viscous_cell.f90:96-109): T = p/(rho r), grad T = grad p/(rho r) - p grad rho/(rho^2 r),
// from the primitive state q, its reciprocal density ir and its gradient record g (rho, p, u..).
template <int DIM, typename R>
__device__ __forceinline__ void side_temp(const Physics& ph, const R* q, R ir, const R* g, R& T, R* gT)
{
    const R k1 = ir * R(ph.irgas);
    T = q[1] * k1;
    const R k2 = T * ir;
#pragma unroll
    for (int e = 0; e < DIM; ++e) gT[e] = k1 * g[DIM + e] - k2 * g[e];
}

// viscous_face for the residual kernel: the same flux from per-side temperatures and
// temperature gradients computed once per cell (side_temp, in K1), velocity gradients gUL/gUR
// (row-major d u_i / d x_j), the face's distance weight dHL (dHR = 1 - dHL) precomputed per
// slot, host-derived constants, and Sutherland's (T/Tref)^1.5 as t sqrt(t) (calc_viscosity,
// EQNS.f90:243-247).  heat = false leaves out the conductive term (flux walls, where the
// imposed flux replaces it).
template <int DIM, typename R>
__device__ __forceinline__ void viscous_face_pre(const Physics& ph, const R* qL, const R* qR, R TL, R TR,
                                                 const R* gTL, const R* gTR, const R* gUL, const R* gUR,
                                                 R dHL, const R* vLR, const R* nf, bool heat, R* flux)
{
    const R dHR = R(1) - dHL;
    R velH[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) velH[e] = dHR * qL[2 + e] + dHL * qR[2 + e];
    const R TH = dHR * TL + dHL * TR;
    R mu;
    if (ph.visc_law == 11) {
        const R t = TH * R(ph.itref);
        mu = adiv(R(ph.visc_dyn) * (t * asqrt(t)) * R(ph.tsum), TH + R(ph.tsuth));
    } else if (ph.visc_law == 14) {
        mu = R(ph.visc_kin) * (dHR * qL[0] + dHL * qR[0]);
    } else {
        mu = R(ph.visc_dyn);
    }
    R G[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) G[i][j] = dHL * gUR[i * DIM + j] + dHR * gUL[i * DIM + j];
    R gavg[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) gavg[e] = dHL * gTR[e] + dHR * gTL[e];
    R gTH;
    if (ph.sch_dis == 10) {
        const R idLR2 = arcp(dotd<DIM>(vLR, vLR));
        R kv[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) kv[e] = idLR2 * vLR[e];
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const R fa = qR[2 + i] - qL[2 + i] - dotd<DIM>(G[i], vLR);
#pragma unroll
            for (int j = 0; j < DIM; ++j) G[i][j] = G[i][j] + fa * kv[j];
        }
        const R pscal = dotd<DIM>(vLR, nf) * idLR2;
        const R fcomp = pscal * (TR - TL);
        R vi[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) vi[e] = nf[e] - pscal * vLR[e];
        gTH = fcomp + dotd<DIM>(gavg, vi);
    } else {
        gTH = dotd<DIM>(gavg, nf);
    }
    R tr = G[0][0];
#pragma unroll
    for (int i = 1; i < DIM; ++i) tr += G[i][i];
    const R addv = -(R(2) / R(3)) * tr;
    R sn[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            R sij = G[i][j] + G[j][i];
            if (i == j) sij += addv;
            s += (mu * sij) * nf[j];
        }
        sn[i] = s;
    }
    const R svn = dotd<DIM>(sn, velH);
#pragma unroll
    for (int e = 0; e < DIM; ++e) flux[2 + e] = flux[2 + e] - sn[e];
    flux[1] = flux[1] - svn;
    if (heat) flux[1] = flux[1] - (mu * R(ph.cp_pr)) * gTH;
}

// ---------------------------------------------------------------------------
// Face Jacobians for the implicit operator: calc_jacflux (calc_jacflux.f90:49-77), the
// Euler flux Jacobian A(n) in (rho, rhoE, rho u..) at a primitive state, weighted by
// calc_jac_hll (calc_jac_hll.f90:46-57) with the flux's wave speeds clipped at zero
// (calc_flux_hllc.f90:185-189).  Row-major NV x NV blocks, dF_i/dU_j.

template <int DIM, typename R>
__device__ __forceinline__ void jacflux(R g, const R* q, const R* n, R vn, R* m)
{
    constexpr int NV = DIM + 2;
    const R g1 = g - R(1), ig1 = R(1) / g1;
    const R v2 = dotd<DIM>(q + 2, q + 2);
    const R en = ig1 * q[1] / q[0] + R(.5) * v2;
    const R* u = q + 2;
    m[0] = R(0);
    m[1] = R(0);
#pragma unroll
    for (int j = 0; j < DIM; ++j) m[2 + j] = n[j];
    m[NV + 0] = vn * (g1 * v2 - g * en);
    m[NV + 1] = vn * g;
#pragma unroll
    for (int j = 0; j < DIM; ++j) m[NV + 2 + j] = (g * en - R(.5) * g1 * v2) * n[j] - g1 * vn * u[j];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        m[(2 + i) * NV + 0] = R(.5) * g1 * v2 * n[i] - vn * u[i];
        m[(2 + i) * NV + 1] = g1 * n[i];
#pragma unroll
        for (int j = 0; j < DIM; ++j) m[(2 + i) * NV + 2 + j] = u[i] * n[j] - (g1 * n[i]) * u[j] + (i == j ? vn : R(0));
    }
}

template <int DIM, typename R>
__device__ __forceinline__ void jac_hll(const Physics& ph, const R* qL, const R* qR, const R* n, R* jL, R* jR)
{
    constexpr int NV = DIM + 2;
    const R g = R(ph.gamma);
    const R vnl = dotd<DIM>(qL + 2, n), vnr = dotd<DIM>(qR + 2, n);
    const R irL = R(1) / qL[0], irR = R(1) / qR[0];
    const R al = sqrt(g * qL[1] * irL), ar = sqrt(g * qR[1] * irR);
    R rrho, rv[DIM], H, a2;
    roe_avg<DIM>(g, R(ph.gig1), qL, qR, irL, irR, rrho, rv, H, a2);
    const R vm = dotd<DIM>(rv, n);
    const R am = sqrt(a2);
    // the flux's own wave speeds; HLLC's for the fluxes outside the HLL family, which the
    // reference does not pair with HLL Jacobians
    R sl, sr;
    hll_speeds((ph.sch_hyp >= 21 && ph.sch_hyp <= 28) ? ph.sch_hyp : 25, g, vnl, vnr, al, ar, vm, am, sl, sr);
    sl = fmin_(sl, R(0));
    sr = fmax_(sr, R(0));
    jacflux<DIM>(g, qL, n, vnl, jL);
    jacflux<DIM>(g, qR, n, vnr, jR);
#pragma unroll
    for (int i = 0; i < NV; ++i) { jL[i * NV + i] -= sl; jR[i * NV + i] -= sr; }
    const R kl = sr / (sr - sl), kr = -sl / (sr - sl);
#pragma unroll
    for (int k = 0; k < NV * NV; ++k) { jL[k] = kl * jL[k]; jR[k] = kr * jR[k]; }
}

// viscous diagonal of the face Jacobians (calc_flux_viscous_cell.f90:149-158)
template <int DIM, typename R>
__device__ __forceinline__ R visc_jac_diag(const Physics& ph, const R* qL, const R* qR, R dHL, const R* d)
{
    const R r = R(ph.rgas);
    const R dHR = R(1) - dHL;
    const R TL = qL[1] / (qL[0] * r), TR = qR[1] / (qR[0] * r);
    const R TH = dHR * TL + dHL * TR;
    const R rhoH = dHR * qL[0] + dHL * qR[0];
    R mu;
    if (ph.visc_law == 11) mu = R(ph.visc_dyn) * dpow(TH / R(ph.tref), R(1.5)) * (R(ph.tref) + R(ph.tsuth)) / (TH + R(ph.tsuth));
    else if (ph.visc_law == 14) mu = R(ph.visc_kin) * rhoH;
    else mu = R(ph.visc_dyn);
    return mu / (dHR * qL[0] + dHL * qR[0]) / sqrt(dotd<DIM>(d, d));
}

// in-register Gauss-Jordan inverse with partial pivoting (FP64)
template <int NV>
__device__ __forceinline__ void invert_block(double* m, double* inv)
{
#pragma unroll
    for (int i = 0; i < NV * NV; ++i) inv[i] = (i % (NV + 1) == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int col = 0; col < NV; ++col) {
        int piv = col;
        double best = fabs(m[col * NV + col]);
#pragma unroll
        for (int r = col + 1; r < NV; ++r) {
            const double v = fabs(m[r * NV + col]);
            if (v > best) { best = v; piv = r; }
        }
#pragma unroll
        for (int r = col; r < NV; ++r) {
            if (r == piv && piv != col) {
#pragma unroll
                for (int k = 0; k < NV; ++k) {
                    double t = m[col * NV + k]; m[col * NV + k] = m[r * NV + k]; m[r * NV + k] = t;
                    t = inv[col * NV + k]; inv[col * NV + k] = inv[r * NV + k]; inv[r * NV + k] = t;
                }
            }
        }
        const double ip = 1.0 / m[col * NV + col];
#pragma unroll
        for (int k = 0; k < NV; ++k) { m[col * NV + k] *= ip; inv[col * NV + k] *= ip; }
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            if (r == col) continue;
            const double f = m[r * NV + col];
#pragma unroll
            for (int k = 0; k < NV; ++k) { m[r * NV + k] -= f * m[col * NV + k]; inv[r * NV + k] -= f * inv[col * NV + k]; }
        }
    }
}

// ---------------------------------------------------------------------------
// K2: fluxes + deterministic gather + fused RK stage update.
// calc_rhs.f90:48-145 (face_state boundary map on the right state), integration_ns_ust.f90,
// flux_to_res.f90:47-142, treelevel_rungekutta.f90:60-105, update_field.f90:38.
// FS: second order, own-side face states come from K1 and the far side's is read
// at its mirror slot; otherwise first order on cell states.  VISC: add the
// viscous flux from cell states and cell gradients of both sides.

// Waves per SIMD the FP64 first-order viscous variant is bounded to: at its natural 170 VGPRs it
// runs 2 waves; 168 VGPRs (3 waves) spill 2 registers and run 12% faster on 16.8M hexes.  The
// other variants keep the compiler's allocation (bounding them one wave higher spills more and
// measured slower).
template <typename R, bool FS, bool VISC, bool JAC>
constexpr int flux_min_waves() { return (VISC && !FS && !JAC && sizeof(R) == 8) ? 3 : 1; }

template <int DIM, typename R, bool FS, bool VISC, int NSM, bool JAC, bool DENSE, bool BND, bool A32>
__device__ __forceinline__ void flux_cell(const KArgs<R>& a, int c);

// Meshes whose cells use every slot run the wavefronts without boundary slots through a copy of
// the cell's code without the boundary branches (as k_gradient).  A32: the FP32 kernels of
// TY_PREC_MIXED_F32SUM, which keep the RK stage-increment sum in a.acc32 (an instance of its own:
// a run-time choice inside one kernel cost the 16.8M-hex NS kernel most of the bytes it saves).
template <int DIM, typename R, bool FS, bool VISC, int NSM, bool JAC, bool DENSE, bool A32>
__global__ __launch_bounds__(BLOCK, (flux_min_waves<R, FS, VISC, JAC>())) void k_flux(KArgs<R> a)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n || (a.stop && *a.stop)) return;
    if (!DENSE || JAC || !flux_interior_path<R, VISC>() || wave_boundary(a.bnd_bits, c))
        flux_cell<DIM, R, FS, VISC, NSM, JAC, DENSE, true, A32>(a, c);
    else
        flux_cell<DIM, R, FS, VISC, NSM, JAC, DENSE, false, A32>(a, c);
}

template <int DIM, typename R, bool FS, bool VISC, int NSM, bool JAC, bool DENSE, bool BND, bool A32>
__device__ __forceinline__ void flux_cell(const KArgs<R>& a, int c)
{
    static_assert(!A32 || sizeof(R) == sizeof(float), "the FP32 increment sum is a form of the FP32 kernels");
    constexpr int NV = DIM + 2;
    const int n = a.n;
    const R gam = R(a.ph.gamma);
    R qc[NV], irc = R(0);
    if (!FS || VISC) load_prim<DIM>(a.us, c, qc, irc);
    double D[JAC ? NV * NV : 1];
    const double kdt = JAC ? *a.d_dt / a.vol[c] : 0.0;
    if (JAC) {
#pragma unroll
        for (int k = 0; k < NV * NV; ++k) D[k] = (k % (NV + 1) == 0) ? 1.0 : 0.0;
    }
    constexpr int VR = DIM * DIM + DIM;
    R gcell[VISC ? VR : 1];
    R Tc = R(0);
    if (VISC) {
        load_rec<VR>(a.grad, c, gcell);
        Tc = qc[1] * (irc * R(a.ph.irgas));
    }
    R res[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) res[v] = R(0);
    // FP32 interior path of the first-order viscous flux: per slot one packed record pb (centre
    // difference, link word) and one pa (area vector, distance weight) in place of six scalar loads.
    // (The FP64 kernel measured 1% slower with the two-plane FP64 records and keeps its scalar loads.)
    constexpr bool PF = !FS && VISC && !JAC && DENSE && !BND && sizeof(R) == 4;
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (!DENSE && k >= a.ns) break;
        decltype(ld_pack(a.g.pb, 0, 0)) pb;
        if (PF) pb = ld_pack(a.g.pb, (size_t)k * n + c, (size_t)a.ns * n);
        const int L = PF ? link_of(pb.w) : a.link[(size_t)k * n + c];
        if (!DENSE && L == SLOT_UNUSED) continue;
        const bool bnd = BND && link_boundary(L);
        const bool other = link_other(L);
        R nf[DIM];
        R A, hpk = R(0);
        if (PF) {
            const auto pa = ld_pack(a.g.pa, (size_t)k * n + c, (size_t)a.ns * n);
            const R S[3] = {R(pa.x), R(pa.y), R(pa.z)};
            A = asqrt(dotd<DIM>(S, S));
            const R ia = arcp(A);
#pragma unroll
            for (int e = 0; e < DIM; ++e) nf[e] = S[e] * ia;
            hpk = R(pa.w);
        } else if (VISC && DENSE) {
            R S[DIM];
            slot_vec_nt<DIM>(a.g.ss, k, n, c, S);
            A = asqrt(dotd<DIM>(S, S));
            const R ia = arcp(A);
#pragma unroll
            for (int e = 0; e < DIM; ++e) nf[e] = S[e] * ia;
        } else {
            A = slot_normal<DIM>(a.g, k, n, c, nf);
        }
        // the far cell's state (first order, viscous flux)
        R qo[NV], iro = R(0);
        if (!FS || VISC) far_state<DIM, BND>(a, k, c, L, qc, qo, iro);
        R qLf[NV], qRf[NV], irL, irR;
        if (FS) {
            const R* own = a.qf + ((size_t)k * n + c) * NV;
            if (!bnd) {
                const R* far = a.qf + ((size_t)link_mirror(L) * n + link_index(L)) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const R ov = own[v], fv = far[v];
                    qLf[v] = other ? fv : ov;
                    qRf[v] = other ? ov : fv;
                }
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) qLf[v] = own[v];
                bc_state<DIM>(a.boco[a.g.bboco[link_index(L)]], nf, qLf, qRf, a.ph, true);
            }
            irL = arcp(qLf[0]);
            irR = arcp(qRf[0]);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                qLf[v] = other ? qo[v] : qc[v];
                qRf[v] = other ? qc[v] : qo[v];
            }
            irL = other ? iro : irc;
            irR = other ? irc : iro;
            if (bnd && a.boco[a.g.bboco[link_index(L)]].kind == BK_WALL_FLUX) {
                // first order still takes the right face state from the face_state map
                // (calc_rhs.f90:103-109), which differs from the ghost cell's on flux walls
                bc_state<DIM>(a.boco[a.g.bboco[link_index(L)]], nf, qc, qRf, a.ph, true);
                irR = arcp(qRf[0]);
            }
        }
        R flux[NV];
        num_flux_pre<DIM>(a.ph.sch_hyp, gam, R(a.ph.ig1), R(a.ph.gig1), qLf, qRf, irL, irR, nf, flux);
        R dk[DIM], dHL = R(0);
        if (VISC) {
            R go[VR];
            if (PF) {
                const R d3[3] = {R(pb.x), R(pb.y), R(pb.z)};
#pragma unroll
                for (int e = 0; e < DIM; ++e) dk[e] = d3[e];
                dHL = hpk;
            } else if (DENSE) {
                slot_vec_nt<DIM>(a.g.sd, k, n, c, dk);
                dHL = __builtin_nontemporal_load(a.g.sh + (size_t)k * n + c);
            } else {
                slot_vec<DIM>(a.g.sd, k, n, c, dk);
                dHL = a.g.sh[(size_t)k * n + c];
            }
            load_rec<VR>(bnd ? a.gbnd : a.grad, link_index(L), go);
            const R To = qo[1] * (iro * R(a.ph.irgas));
            // ns_bocoflux (ns_bocoflux.f90:47-111): on flux and adiabatic walls the imposed heat
            // flux replaces the conductive one
            const bool fwall = bnd && a.boco[a.g.bboco[link_index(L)]].kind == BK_WALL_FLUX;
            viscous_face_pre<DIM>(a.ph, other ? qo : qc, other ? qc : qo, other ? To : Tc, other ? Tc : To,
                                  (other ? go : gcell) + DIM * DIM, (other ? gcell : go) + DIM * DIM,
                                  other ? go : gcell, other ? gcell : go, dHL, dk, nf, !fwall, flux);
            if (fwall) flux[1] = flux[1] + R(a.boco[a.g.bboco[link_index(L)]].param[0]);
        }
        if (JAC) {
            // build_implicit_bdlu.f90:44-83: diag(ic1) += jacL.A dt/V, diag(ic2) -= jacR.A dt/V
            R jL[NV * NV], jR[NV * NV];
            jac_hll<DIM>(a.ph, qLf, qRf, nf, jL, jR);
            if (VISC) {
                const R id = visc_jac_diag<DIM>(a.ph, other ? qo : qc, other ? qc : qo, dHL, dk);
#pragma unroll
                for (int i = 0; i < NV; ++i) { jR[i * NV + i] -= id; jL[i * NV + i] += id; }
            }
            if (a.bdlu_B && bnd) {
                // ns_bocojacobian.f90:44-92: slip wall jL += jR dQR/dQL (mirror of the momentum
                // block), supersonic outlet jL += jR; then jR = 0
                const int kind = a.boco[a.g.bboco[link_index(L)]].kind;
                if (kind == BK_SYM || kind == BK_OUTLET_SUP) {
                    R Sm[NV * NV];
#pragma unroll
                    for (int k2 = 0; k2 < NV * NV; ++k2) Sm[k2] = (k2 % (NV + 1) == 0) ? R(1) : R(0);
                    if (kind == BK_SYM) {
#pragma unroll
                        for (int i = 0; i < DIM; ++i)
#pragma unroll
                            for (int j = 0; j < DIM; ++j) Sm[(2 + i) * NV + 2 + j] -= R(2) * (nf[i] * nf[j]);
                    }
#pragma unroll
                    for (int i = 0; i < NV; ++i)
#pragma unroll
                        for (int j = 0; j < NV; ++j) {
                            R sacc = R(0);
#pragma unroll
                            for (int l = 0; l < NV; ++l) sacc += jR[i * NV + l] * Sm[l * NV + j];
                            jL[i * NV + j] += sacc;
                        }
#pragma unroll
                    for (int k2 = 0; k2 < NV * NV; ++k2) jR[k2] = R(0);
                }
            }
            if (other) {
#pragma unroll
                for (int k2 = 0; k2 < NV * NV; ++k2) D[k2] -= (double)(A * jR[k2]) * kdt;
            } else {
#pragma unroll
                for (int k2 = 0; k2 < NV * NV; ++k2) D[k2] += (double)(A * jL[k2]) * kdt;
            }
            if (a.bdlu_B) {
                // off-diagonal block of this row: upper(f) = jacR.A dt/V_owner on the owner's row,
                // lower(f) = -jacL.A dt/V_other on the other cell's row; a periodic face is two
                // boundary faces in the reference, each coupling its cell to its own ghost
                const unsigned fwk = (unsigned)a.fw[(size_t)k * n + c];
                const int pj = (int)(fwk & FACE_MASK) - a.per_base;
                const bool per = (fwk & PERIODIC_BIT) != 0u;
                double* Bk = a.bdlu_B + (size_t)k * NV * NV * n + c;
                int tgt;
                if (other) {
#pragma unroll
                    for (int k2 = 0; k2 < NV * NV; ++k2) Bk[(size_t)k2 * n] = -((double)(A * jL[k2]) * kdt);
                    tgt = per ? n + a.nb_dev + 2 * pj + 1 : link_index(L);
                } else {
#pragma unroll
                    for (int k2 = 0; k2 < NV * NV; ++k2) Bk[(size_t)k2 * n] = (double)(A * jR[k2]) * kdt;
                    tgt = bnd ? n + link_index(L) : (per ? n + a.nb_dev + 2 * pj : link_index(L));
                }
                a.bdlu_tgt[(size_t)k * n + c] = tgt;
            }
        }
        if (JAC && a.bdlu_grhs) {
            // ghost rows of the residual: the boundary-flux sums flux_to_res leaves there
            // (flux_to_res.f90:90-113, ghost = ic2 of its boundary face)
            const unsigned fwk = (unsigned)a.fw[(size_t)k * n + c];
            const bool per = (fwk & PERIODIC_BIT) != 0u;
            if (bnd || per) {
                const int g = bnd ? link_index(L) : a.nb_dev + 2 * ((int)(fwk & FACE_MASK) - a.per_base) + (other ? 1 : 0);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const double fa = (double)(A * flux[v]);
                    a.bdlu_grhs[(size_t)v * a.bdlu_ld + g] = other ? -fa : fa;
                }
            }
        }
        // flux_to_res.f90:90-113: R(ic1) -= F.A, R(ic2) += F.A
        if (other) {
#pragma unroll
            for (int v = 0; v < NV; ++v) res[v] = res[v] + A * flux[v];
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) res[v] = res[v] - A * flux[v];
        }
    }
    // flux_to_res.f90:126-142: R = dt/V * sum (the cell's own time step with loc_stab_cond)
    const double dt = a.dtloc ? a.dtloc[c] : *a.d_dt;
    const double coef = dt / a.vol[c];
    double Rv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Rv[v] = coef * (double)res[v];
    if (a.mode == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) a.res_out[(size_t)v * n + c] = Rv[v];
        if (JAC) {
            if (a.diag_out) {
#pragma unroll
                for (int k = 0; k < NV * NV; ++k) a.diag_out[(size_t)k * n + c] = D[k];
            }
            if (a.dinv_out) {
                double inv[NV * NV];
                invert_block<NV>(D, inv);
#pragma unroll
                for (int k = 0; k < NV * NV; ++k) a.dinv_out[(size_t)k * n + c] = (R)inv[k];
            }
        }
        return;
    }
    // RK bookkeeping (treelevel_rungekutta.f90:60-105).  The increments are summed in FP64 and
    // kept between stages in FP64, or in FP32 in a.acc32 (A32, TY_PREC_MIXED_F32SUM: each R_j comes
    // from FP32 face fluxes, whose rounding, relative to the sum of |face flux|, already exceeds
    // the FP32 rounding of the partial sum); the state stays FP64 (U = U0 + sum at the last stage).
    constexpr bool a32 = A32;
    double acc[NV];
    if (a.stage == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = Rv[v] * a.b;
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            acc[v] = (a32 ? (double)a.acc32[(size_t)v * n + c] : a.acc[(size_t)v * n + c]) + a.b * Rv[v];
    }
    if (a.store_slot >= 0) {
        double* rs = a.rstore + (size_t)a.store_slot * NV * n;
#pragma unroll
        for (int v = 0; v < NV; ++v) rs[(size_t)v * n + c] = Rv[v];
    }
    // the next stage state in FP64, stored as its primitive record (pressure recovered in FP64)
    double un[NV];
    if (a.stage < a.nstage) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (a32)
                a.acc32[(size_t)v * n + c] = (float)acc[v];
            else
                a.acc[(size_t)v * n + c] = acc[v];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double u = a.u0[(size_t)v * n + c];
            for (int j = 0; j < a.stage - 1; ++j) {
                const int sl = a.read_slot[j];
                if (sl >= 0) u = u + a.a[j] * a.rstore[((size_t)sl * NV + v) * n + c];
            }
            un[v] = u + a.a[a.stage - 1] * Rv[v];
        }
    } else {
        // update_field.f90:38: U = U0 + residu
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            un[v] = a.u0[(size_t)v * n + c] + acc[v];
            a.u0[(size_t)v * n + c] = un[v];
        }
        if (a.rnorm) {
            // update_field.f90:42-60: |residu| / refsca for rho and rhoE, the momentum vector's norm
            // / refvec
            double m2 = acc[2] * acc[2];
#pragma unroll
            for (int d = 1; d < DIM; ++d) m2 += acc[2 + d] * acc[2 + d];
            a.rnorm[c] = fabs(acc[0]) / a.rref[0] + fabs(acc[1]) / a.rref[1] + sqrt(m2) / a.rref[2];
        }
    }
    store_prim_n<DIM>(un, a.ph.gamma - 1.0, a.us_next + (size_t)c * NV);
}

// ---------------------------------------------------------------------------
// time step: calc_ns_timestep.f90:52-102, global min (calc_zonetimestep.f90:95)

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin_(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_dt_local(DtArgs a)
{
    constexpr int NV = DIM + 2;
    __shared__ double smin[BLOCK / 64];
    __shared__ int sbad[BLOCK / 64];
    const int lb = xcd_block();
    const int c = lb * BLOCK + threadIdx.x;
    if (a.stop && *a.stop) return;  // uniform over the grid
    double dt = HUGE_VAL;
    int bad = 0;
    if (c < a.n) {
        const int n = a.n;
        const double gg1 = a.gamma * (a.gamma - 1.0);
        const double rho = a.u0[c];
        const double irho = 1.0 / rho;
        double rv2 = a.u0[(size_t)2 * n + c] * a.u0[(size_t)2 * n + c];
#pragma unroll
        for (int d = 1; d < DIM; ++d) rv2 += a.u0[(size_t)(2 + d) * n + c] * a.u0[(size_t)(2 + d) * n + c];
        const double a2 = (a.u0[(size_t)n + c] - .5 * rv2 * irho) * gg1 * irho;
        if (!(a2 > 0.0)) bad = 1;
        dt = *a.cfl * 2.0 * a.vol[c] / a.sumA[c];
        dt = dt / (sqrt(rv2) * irho + sqrt(a2));
        if (!(dt == dt)) bad = 1;
        if (a.dtloc) a.dtloc[c] = dt;
        (void)NV;
    }
    dt = wave_min(dt);
    bad = wave_sum(bad);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { smin[wid] = dt; sbad[wid] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = smin[0];
        int b = sbad[0];
        for (int i = 1; i < BLOCK / 64; ++i) { m = fmin_(m, smin[i]); b += sbad[i]; }
        a.part[lb] = m;
        a.bad[lb] = b;
    }
}

static __global__ __launch_bounds__(1024) void k_dt_final(const double* part, const int* badp, int nb, double dtmax,
                                                   double* d_dt, int* d_err)
{
    __shared__ double smin[16];
    __shared__ int sbad[16];
    double m = HUGE_VAL;
    int b = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) { m = fmin_(m, part[i]); b += badp[i]; }
    m = wave_min(m);
    b = wave_sum(b);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { smin[wid] = m; sbad[wid] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = smin[0];
        int bb = sbad[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { r = fmin_(r, smin[i]); bb += sbad[i]; }
        r = fmin_(r, dtmax);
        *d_dt = r;
        if (bb > 0) *d_err = 1;
    }
}

// stage-state (primitive) records from a [nv][n] FP64 conservative state
template <typename R>
__global__ void k_to_stage(const double* __restrict__ src, R* __restrict__ dst, int nv, int n, double g1)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double u[NVMAX];
    for (int v = 0; v < nv; ++v) u[v] = src[(size_t)v * n + c];
    store_prim<R>(u, nv, g1, dst + (size_t)c * nv);
}

// MONOTONIC1 on interior and periodic faces: each cell limits the pair (own, far) of K1's face
// states of every slot with the two cell values, in the face's canonical orientation, and keeps its
// own member; both cells of a face compute the same pair.  Boundary slots were limited by K1.
template <int DIM, typename R, int NSM, bool DENSE>
__global__ __launch_bounds__(BLOCK) void k_mono1(KArgs<R> a)
{
    constexpr int NV = DIM + 2;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n || (a.stop && *a.stop)) return;
    const int n = a.n;
    R qc[NV];
    load_prim<DIM>(a.us, c, qc);
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (!DENSE && k >= a.ns) break;
        const int L = a.link[(size_t)k * n + c];
        if (!DENSE && L == SLOT_UNUSED) continue;
        const R* own = a.qf_out + ((size_t)k * n + c) * NV;
        R* out = a.qf_mono + ((size_t)k * n + c) * NV;
        if (link_boundary(L)) {
#pragma unroll
            for (int v = 0; v < NV; ++v) out[v] = own[v];
            continue;
        }
        const R* far = a.qf_out + ((size_t)link_mirror(L) * n + link_index(L)) * NV;
        R qo[NV];
        load_prim<DIM>(a.us, link_index(L), qo);
        const bool other = link_other(L);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const R x1 = other ? far[v] : own[v], x2 = other ? own[v] : far[v];
            const R qa = other ? qo[v] : qc[v], qb = other ? qc[v] : qo[v];
            out[v] = other ? mono_right(11, x1, x2, qa, qb) : mono_left(11, x1, x2, qa, qb);
        }
    }
}

// ---------------------------------------------------------------------------
// launchers

// gradient_variant, rec_variant and flux_variant are instantiated one per translation unit by the
// build (they are the expensive part to compile); launch_gradient/launch_flux only pick one.
template <int DIM, typename R, int NSM, bool DENSE, bool GOUT>
void rec_variant(const KArgs<R>& a, hipStream_t s)
{
    const dim3 grid(grid_blocks(a.n)), block(BLOCK);
    if (a.ph.limiter == 'W')
        hipLaunchKernelGGL((k_gradient<DIM, R, NSM, REC_VENKAT, DENSE, GOUT>), grid, block, 0, s, a);
    else if (a.ph.barth)
        hipLaunchKernelGGL((k_gradient<DIM, R, NSM, REC_MUSCL_BARTH, DENSE, GOUT>), grid, block, 0, s, a);
    else if (a.ph.mono)
        hipLaunchKernelGGL((k_gradient<DIM, R, NSM, REC_MUSCL_MONO, DENSE, GOUT>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_gradient<DIM, R, NSM, REC_MUSCL, DENSE, GOUT>), grid, block, 0, s, a);
}

template <int DIM, typename R, int NSM, bool DENSE>
void gradient_variant(const KArgs<R>& a, hipStream_t s)
{
    const dim3 grid(grid_blocks(a.n)), block(BLOCK);
    if (!a.qf_out) hipLaunchKernelGGL((k_gradient<DIM, R, NSM, REC_NONE, DENSE, true>), grid, block, 0, s, a);
    else if (a.grad_out) rec_variant<DIM, R, NSM, DENSE, true>(a, s);
    else rec_variant<DIM, R, NSM, DENSE, false>(a, s);
    if (a.qf_out && a.qf_mono) hipLaunchKernelGGL((k_mono1<DIM, R, NSM, DENSE>), grid, block, 0, s, a);
}

// slot-count variant: dense 4 (tets, quads) or 6 (hexes) slots per cell, otherwise guarded loops
template <int DIM, typename R>
hipError_t launch_gradient(const KArgs<R>& a, hipStream_t s)
{
    if (a.dense && a.ns == 4) gradient_variant<DIM, R, 4, true>(a, s);
    else if (a.dense && a.ns == 6) gradient_variant<DIM, R, 6, true>(a, s);
    else if (a.ns <= 4) gradient_variant<DIM, R, 4, false>(a, s);
    else gradient_variant<DIM, R, MAX_SLOTS, false>(a, s);
    return hipGetLastError();
}

template <int DIM, typename R, int NSM, bool JAC, bool DENSE, bool A32>
void flux_launch(const KArgs<R>& a, hipStream_t s)
{
    const dim3 grid(grid_blocks(a.n)), block(BLOCK);
    const bool fs = a.ph.method != 'N';
    const bool visc = a.ph.viscous != 0;
    if (fs && visc) hipLaunchKernelGGL((k_flux<DIM, R, true, true, NSM, JAC, DENSE, A32>), grid, block, 0, s, a);
    else if (fs) hipLaunchKernelGGL((k_flux<DIM, R, true, false, NSM, JAC, DENSE, A32>), grid, block, 0, s, a);
    else if (visc) hipLaunchKernelGGL((k_flux<DIM, R, false, true, NSM, JAC, DENSE, A32>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_flux<DIM, R, false, false, NSM, JAC, DENSE, A32>), grid, block, 0, s, a);
}

template <int DIM, typename R, int NSM, bool JAC, bool DENSE>
void flux_variant(const KArgs<R>& a, hipStream_t s)
{
    if constexpr (sizeof(R) == sizeof(float) && !JAC) {
        if (a.acc32 && a.mode != 1) {
            flux_launch<DIM, R, NSM, JAC, DENSE, true>(a, s);
            return;
        }
    }
    flux_launch<DIM, R, NSM, JAC, DENSE, false>(a, s);
}

template <int DIM, typename R>
hipError_t launch_flux(const KArgs<R>& a_in, hipStream_t s)
{
    KArgs<R> a = a_in;
    if (a.qf_mono) a.qf = a.qf_mono;  // MONOTONIC1: the pair-limited face states
    const bool jac = a.dinv_out != nullptr || a.diag_out != nullptr;
    if (jac) {  // implicit diagonal blocks: guarded loops only
        if (a.ns <= 4) flux_variant<DIM, R, 4, true, false>(a, s);
        else flux_variant<DIM, R, MAX_SLOTS, true, false>(a, s);
    } else if (a.dense && a.ns == 4) {
        flux_variant<DIM, R, 4, false, true>(a, s);
    } else if (a.dense && a.ns == 6) {
        flux_variant<DIM, R, 6, false, true>(a, s);
    } else if (a.ns <= 4) {
        flux_variant<DIM, R, 4, false, false>(a, s);
    } else {
        flux_variant<DIM, R, MAX_SLOTS, false, false>(a, s);
    }
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_timestep(const DtArgs& a, double dtmax, double* d_dt, int* d_err, hipStream_t s)
{
    const int nb = grid_blocks(a.n);
    hipLaunchKernelGGL((k_dt_local<DIM>), dim3(nb), dim3(BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1024), 0, s, (const double*)a.part, (const int*)a.bad, nb, dtmax,
                       d_dt, d_err);
    return hipGetLastError();
}
template <typename R>
hipError_t launch_to_stage(const double* src, R* dst, int nv, int n, double gamma, hipStream_t s)
{
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL((k_to_stage<R>), dim3(nb), dim3(256), 0, s, src, dst, nv, n, gamma - 1.0);
    return hipGetLastError();
}

}  // namespace ty

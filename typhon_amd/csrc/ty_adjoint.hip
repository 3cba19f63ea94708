// Tangent-linear and adjoint of the first-order residual R(U; S, V, q_in) =
// -dt/V sum_f F(U_L, U_R; n) A (calc_rhs.f90:10, flux_to_res.f90:47-142) at a
// fixed time step, in FP64.  Inputs of the derivative: the conservative state of
// the interior cells, the area vector S = A n of every face, the cell volumes and
// the ghost primitive state of the supersonic inlets.
//
// The tangent keeps the primal layout: one work-item per cell walking its face
// slots, every face flux evaluated in the face's canonical orientation with dual
// numbers (value + one derivative), a deterministic per-cell sum.  The adjoint
// contracts the derivative of each face flux with the face's adjoint weight
// m = sigma A (lambda_c dt/V_c - lambda_nb dt/V_nb).  For HLLC/HLLCB it runs one
// work-item per device face through the hand-written reverse sweep of the flux
// (one forward + one reverse pass gives the derivative w.r.t. both states and the
// normal; boundary ghost states are chained through their own Jacobian), writes
// one record per face side, and a per-cell kernel sums the records in slot order:
// (dR/dU)^T lambda without atomics.  The other fluxes take one dual-number pass
// per input direction inside a per-cell kernel.  Every output has one writer.
#include "ty_kernels_impl.h"
#include "ty_dual.h"
#include "ty_hllc_rev.h"
#include "ty_visc_rev.h"
#include "ty_adjoint.h"

namespace ty {

// (dq/dU)^T qb added to Ub, from the primitive state q (rho, p, u..): prim_rev of ty_hllc_rev.h with
// the velocity taken from the record
template <int DIM>
__device__ __forceinline__ void prim_rev_q(const double* q, double g1, const double* qb, double* Ub)
{
    const double ir = 1.0 / q[0];
    double uu = 0.0, ub_u = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        uu += q[2 + d] * q[2 + d];
        ub_u += qb[2 + d] * q[2 + d];
    }
    Ub[0] += qb[0] + qb[1] * g1 * 0.5 * uu - ub_u * ir;
    Ub[1] += qb[1] * g1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) Ub[2 + d] += -g1 * q[2 + d] * qb[1] + qb[2 + d] * ir;
}

template <int DIM>
__device__ __forceinline__ void load_q(const AdArgs& t, int c, double* q)
{
    const double* r = t.qrec + (size_t)c * (DIM + 2);
#pragma unroll
    for (int v = 0; v < DIM + 2; ++v) q[v] = r[v];
}

template <int DIM, typename R>
__device__ __forceinline__ void cons_to_prim(const R* u, R g1, R* q)
{
    // calc_varprim_ns (SOURCE/EQNS/calc_varprim_ns.f90:47-52), same operation order as load_prim
    const R rho = u[0];
    R vel[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) vel[d] = u[2 + d] / rho;
    R sq = vel[0] * vel[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) sq += vel[d] * vel[d];
    const R ec = R(.5) * rho * sq;
    q[0] = rho;
    q[1] = g1 * (u[1] - ec);
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[2 + d] = vel[d];
}

// Flux of one face slot in dual arithmetic.  Uc: this cell's conservative state;
// Uo: the far cell's (ignored on boundary slots); nf: unit normal (owner -> other);
// dqg: derivative of a supersonic-inlet ghost state (primitive, or null).
template <int DIM>
__device__ __forceinline__ void slot_flux(const KArgs<double>& a, const SlotInfo& s, const Dual* Uc, const Dual* Uo,
                                          const Dual* nf, const double* dqg, int seed_ghost, Dual* F)
{
    constexpr int NV = DIM + 2;
    const Dual g1(a.ph.gamma - 1.0);
    Dual qc[NV], qL[NV], qR[NV];
    cons_to_prim<DIM>(Uc, g1, qc);
    if (s.nb >= 0) {
        Dual qo[NV];
        cons_to_prim<DIM>(Uo, g1, qo);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            qL[v] = s.other ? qo[v] : qc[v];
            qR[v] = s.other ? qc[v] : qo[v];
        }
    } else {
        const BocoDev bc = a.boco[a.g.bboco[-s.nb - 1]];
#pragma unroll
        for (int v = 0; v < NV; ++v) qL[v] = qc[v];
        bc_state<DIM>(bc, nf, qL, qR, a.ph);
        if (bc.kind == BK_INLET_SUP) {
            if (dqg) {
#pragma unroll
                for (int v = 0; v < NV; ++v) qR[v].d += dqg[v];
            }
#pragma unroll
            for (int v = 0; v < NV; ++v)
                if (v == seed_ghost) qR[v].d = 1.0;
        }
    }
    num_flux<DIM>(a.ph.sch_hyp, Dual(a.ph.gamma), qL, qR, nf, F);
}

// A face's area-vector sensitivity from gn = (dF/dn)^T m and the flux (written by the owner):
// lambda.R depends on S = A n through n = S/|S| and A = |S|.
template <int DIM>
__device__ __forceinline__ void face_sens(const AdArgs& t, int f, const double* nv, double A, const double* mv,
                                          const double* Fv, const double* gn)
{
    constexpr int NV = DIM + 2;
    double ndg = 0.0, ga = 0.0;
#pragma unroll
    for (int e = 0; e < DIM; ++e) ndg += nv[e] * gn[e];
#pragma unroll
    for (int v = 0; v < NV; ++v) ga += (mv[v] / A) * Fv[v];
#pragma unroll
    for (int e = 0; e < 3; ++e) t.gs[(size_t)e * t.nf + f] = e < DIM ? (gn[e] - nv[e] * ndg) / A + nv[e] * ga : 0.0;
}

template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_res_tangent(AdArgs t)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    Dual Uc[NV], res[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        Uc[v] = Dual(a.us[(size_t)v * n + c], t.du ? t.du[(size_t)v * n + c] : 0.0);
        res[v] = Dual(0.0);
    }
    for (int k = 0; k < a.ns; ++k) {
        const int w = a.fw[(size_t)k * n + c];
        if (w == SLOT_UNUSED) continue;
        const SlotInfo s = decode(w, a.nbr[(size_t)k * n + c]);
        double nv[DIM];
        face_n<DIM>(a.g, s.f, nv);
        const double A = a.g.fs[s.f];
        // S = A n: dA = n.dS, dn = (dS - n dA) / A
        Dual nd[DIM], Ad(A);
        double dS[DIM], dA = 0.0;
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            dS[e] = t.ds ? t.ds[(size_t)e * t.nf + s.f] : 0.0;
            dA += nv[e] * dS[e];
        }
        Ad.d = dA;
#pragma unroll
        for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e], (dS[e] - nv[e] * dA) / A);
        Dual Uo[NV];
        if (s.nb >= 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                Uo[v] = Dual(a.us[(size_t)v * n + s.nb], t.du ? t.du[(size_t)v * n + s.nb] : 0.0);
        }
        const double* dqg = nullptr;
        if (s.nb < 0 && t.dq) dqg = t.dq + (size_t)a.g.bboco[-s.nb - 1] * NVMAX;
        Dual F[NV];
        slot_flux<DIM>(a, s, Uc, Uo, nd, dqg, -1, F);
        // flux_to_res.f90:90-113
#pragma unroll
        for (int v = 0; v < NV; ++v) res[v] = s.other ? res[v] + Ad * F[v] : res[v] - Ad * F[v];
    }
    const Dual Vd(a.vol[c], t.dvol ? t.dvol[c] : 0.0);
    const Dual coef = Dual(*a.d_dt) / Vd;
#pragma unroll
    for (int v = 0; v < NV; ++v) t.out[(size_t)v * n + c] = (coef * res[v]).d;
}

// Adjoint contributions of one face slot of cell c by dual-number passes (any flux, any
// boundary map): gu += (dF/dU_c)^T m, the face's sensitivities if c owns it, and the flux.
template <int DIM>
__device__ __forceinline__ void adj_slot_dual(const AdArgs& t, int c, const SlotInfo& s, const double* Uc,
                                              const double* Uon, const double* nv, double A, const double* mv,
                                              double* gu, double* Fv)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    Dual Uo[NV], Ud[NV], nd[DIM], F[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Uo[v] = Dual(Uon[v]);
#pragma unroll
    for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e]);
    for (int j = 0; j < NV; ++j) {
#pragma unroll
        for (int v = 0; v < NV; ++v) Ud[v] = Dual(Uc[v], v == j ? 1.0 : 0.0);
        slot_flux<DIM>(a, s, Ud, Uo, nd, nullptr, -1, F);
        double acc = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) acc += mv[v] * F[v].d;
        gu[j] += acc;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        Fv[v] = F[v].v;
        Ud[v] = Dual(Uc[v]);
    }
    if (!s.other && t.gs) {
        double gn[DIM];
        for (int e = 0; e < DIM; ++e) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) nd[i] = Dual(nv[i], i == e ? 1.0 : 0.0);
            slot_flux<DIM>(a, s, Ud, Uo, nd, nullptr, -1, F);
            double acc = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc += mv[v] * F[v].d;
            gn[e] = acc;
        }
#pragma unroll
        for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e]);
        face_sens<DIM>(t, s.f, nv, A, mv, Fv, gn);
    }
    if (s.nb < 0 && t.gq) {
        const int bi = -s.nb - 1;
        const bool inlet = a.boco[a.g.bboco[bi]].kind == BK_INLET_SUP;
        for (int j = 0; j < NV; ++j) {
            double acc = 0.0;
            if (inlet) {
                slot_flux<DIM>(a, s, Ud, Uo, nd, nullptr, j, F);
#pragma unroll
                for (int v = 0; v < NV; ++v) acc += mv[v] * F[v].d;
            }
            t.gq[(size_t)j * t.nb + bi] = acc;
        }
    }
}

// adjoint weight of slot k's flux: d(lambda.R)/dF = sigma A (lambda_c dt/V_c - lambda_nb dt/V_nb)
template <int DIM>
__device__ __forceinline__ void adj_weight(const AdArgs& t, const SlotInfo& s, double A, const double* lc, double dt,
                                           double* Uo, double* mv)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int n = a.n;
    const double sg = s.other ? 1.0 : -1.0;
    if (s.nb >= 0) {
        const double dto = dt / a.vol[s.nb];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Uo[v] = a.us[(size_t)v * n + s.nb];
            mv[v] = sg * A * (lc[v] - t.lam[(size_t)v * n + s.nb] * dto);
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Uo[v] = 0.0;
            mv[v] = sg * A * lc[v];
        }
    }
}

// Every slot by dual-number passes (the fluxes without a hand-written reverse sweep).
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_res_adjoint(AdArgs t)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    const double dt = *a.d_dt;
    const double dtv = dt / a.vol[c];
    double Uc[NV], lc[NV], gu[NV], res[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        Uc[v] = a.us[(size_t)v * n + c];
        lc[v] = t.lam[(size_t)v * n + c] * dtv;
        gu[v] = 0.0;
        res[v] = 0.0;
    }
    for (int k = 0; k < a.ns; ++k) {
        const int w = a.fw[(size_t)k * n + c];
        if (w == SLOT_UNUSED) continue;
        const SlotInfo s = decode(w, a.nbr[(size_t)k * n + c]);
        double nv[DIM];
        face_n<DIM>(a.g, s.f, nv);
        const double A = a.g.fs[s.f];
        double Uo[NV], mv[NV], Fv[NV];
        adj_weight<DIM>(t, s, A, lc, dt, Uo, mv);
        adj_slot_dual<DIM>(t, c, s, Uc, Uo, nv, A, mv, gu, Fv);
        const double sg = s.other ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) res[v] += sg * A * Fv[v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) t.gu[(size_t)v * n + c] = gu[v];
    if (t.gvol) {
        // d(lambda.R_c)/dV_c = -lambda_c . R_c / V_c with R_c = dt/V_c res_c
        double acc = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) acc += lc[v] * res[v];
        t.gvol[c] = -acc / a.vol[c];
    }
}

// ---------------------------------------------------------------------------
// Viscous first-order residual: the flux also reads both cells' gradients of the primitive
// state (LSQ, g_c = sum_k w_k (q_k - q_c), ghost states across boundary slots), so the
// derivative reaches the neighbours' neighbours through the gradients.

// tangent of the cell gradients (value in k.grad): dg_c = sum_k w_k (dq_k - dq_c)
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_grad_tangent(AdArgs t)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    const Dual g1(a.ph.gamma - 1.0);
    Dual Uc[NV], qc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) Uc[v] = Dual(a.us[(size_t)v * n + c], t.du ? t.du[(size_t)v * n + c] : 0.0);
    cons_to_prim<DIM>(Uc, g1, qc);
    double dg[NV * DIM];
#pragma unroll
    for (int k = 0; k < NV * DIM; ++k) dg[k] = 0.0;
    for (int k = 0; k < a.ns; ++k) {
        const int L = a.link[(size_t)k * n + c];
        if (L == SLOT_UNUSED) continue;
        Dual qo[NV];
        if (!link_boundary(L)) {
            const int nb = link_index(L);
            Dual Uo[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Uo[v] = Dual(a.us[(size_t)v * n + nb], t.du ? t.du[(size_t)v * n + nb] : 0.0);
            cons_to_prim<DIM>(Uo, g1, qo);
        } else {
            const int bi = link_index(L);
            double nv[DIM];
            slot_normal<DIM>(a.g, k, n, c, nv);
            Dual nd[DIM];
#pragma unroll
            for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e]);
            const BocoDev bc = a.boco[a.g.bboco[bi]];
            bc_state<DIM>(bc, nd, qc, qo, a.ph);
            if (bc.kind == BK_INLET_SUP && t.dq) {
#pragma unroll
                for (int v = 0; v < NV; ++v) qo[v].d += t.dq[(size_t)a.g.bboco[bi] * NVMAX + v];
            }
        }
        double wk[DIM];
        slot_vec<DIM>(a.g.sw, k, n, c, wk);
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int e = 0; e < DIM; ++e) dg[v * DIM + e] += wk[e] * (qo[v].d - qc[v].d);
    }
    double* out = t.dgrad + (size_t)c * NV * DIM;
#pragma unroll
    for (int k = 0; k < NV * DIM; ++k) out[k] = dg[k];
}

// tangent of the viscous first-order residual (state, volume and inlet directions)
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_res_tangent_visc(AdArgs t)
{
    constexpr int NV = DIM + 2, NG = NV * DIM;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    const Dual g1(a.ph.gamma - 1.0);
    Dual Uc[NV], qc[NV], gc[NG], res[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        Uc[v] = Dual(a.us[(size_t)v * n + c], t.du ? t.du[(size_t)v * n + c] : 0.0);
        res[v] = Dual(0.0);
    }
    cons_to_prim<DIM>(Uc, g1, qc);
#pragma unroll
    for (int k = 0; k < NG; ++k) gc[k] = Dual(a.grad[(size_t)c * NG + k], t.dgrad[(size_t)c * NG + k]);
    for (int k = 0; k < a.ns; ++k) {
        const int L = a.link[(size_t)k * n + c];
        if (L == SLOT_UNUSED) continue;
        const bool other = link_other(L);
        double nv[DIM];
        const double A = slot_normal<DIM>(a.g, k, n, c, nv);
        Dual nd[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e]);
        Dual qo[NV], go[NG];
        if (!link_boundary(L)) {
            const int nb = link_index(L);
            Dual Uo[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) Uo[v] = Dual(a.us[(size_t)v * n + nb], t.du ? t.du[(size_t)v * n + nb] : 0.0);
            cons_to_prim<DIM>(Uo, g1, qo);
#pragma unroll
            for (int q = 0; q < NG; ++q) go[q] = Dual(a.grad[(size_t)nb * NG + q], t.dgrad[(size_t)nb * NG + q]);
        } else {
            const int bi = link_index(L);
            const BocoDev bc = a.boco[a.g.bboco[bi]];
            bc_state<DIM>(bc, nd, qc, qo, a.ph);
            if (bc.kind == BK_INLET_SUP && t.dq) {
#pragma unroll
                for (int v = 0; v < NV; ++v) qo[v].d += t.dq[(size_t)a.g.bboco[bi] * NVMAX + v];
            }
            bc_grad<DIM>(bc, nd, gc, go);
        }
        const Dual* qL = other ? qo : qc;
        const Dual* qR = other ? qc : qo;
        Dual F[NV];
        num_flux<DIM>(a.ph.sch_hyp, Dual(a.ph.gamma), qL, qR, nd, F);
        double dk[DIM];
        slot_vec<DIM>(a.g.sd, k, n, c, dk);
        Dual dd[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) dd[e] = Dual(dk[e]);
        viscous_face<DIM>(a.ph, qL, qR, a.g.sh[(size_t)k * n + c], dd, nd, other ? go : gc, other ? gc : go, F);
        const Dual Ad(A);
#pragma unroll
        for (int v = 0; v < NV; ++v) res[v] = other ? res[v] + Ad * F[v] : res[v] - Ad * F[v];
    }
    const Dual Vd(a.vol[c], t.dvol ? t.dvol[c] : 0.0);
    const Dual coef = Dual(*a.d_dt) / Vd;
#pragma unroll
    for (int v = 0; v < NV; ++v) t.out[(size_t)v * n + c] = (coef * res[v]).d;
}

// per device face: the viscous distance weight dHL and the canonical centre difference
// d = x_other - x_owner (the ghost centre on boundary faces), from the owner's slot geometry
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_adj_face_geom(AdArgs t)
{
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    for (int k = 0; k < a.ns; ++k) {
        const int w = a.fw[(size_t)k * n + c];
        if (w == SLOT_UNUSED || w < 0) continue;  // unused, or this cell is the far side
        const int f = (int)((unsigned)w & FACE_MASK);
        double dk[DIM];
        slot_vec<DIM>(a.g.sd, k, n, c, dk);
        double* g = const_cast<double*>(t.fgeo);
        g[f] = a.g.sh[(size_t)k * n + c];
#pragma unroll
        for (int e = 0; e < DIM; ++e) g[(size_t)(1 + e) * t.nf + f] = dk[e];
    }
}



// HLLC/HLLCB faces, one work-item per device face: one forward and one reverse sweep of the face
// flux (the cell-parallel form would evaluate every two-sided face twice).  The face weight is
// m = A (lambda_other dt/V_other - lambda_owner dt/V_owner), and m = -A lambda_owner dt/V_owner on a
// boundary face, whose ghost state qR = B(qL, n) (and, viscous, ghost gradient gR = B_g(gL, n)) is
// chained through the transpose of B's Jacobian (dual-number passes of the boundary template,
// cheap next to the flux).  Cell data come as one record per cell (primitive state, lambda dt/V).
// Each face side gets a record [(dF/dU_side)^T m, (viscous) the adjoints of the side's velocity
// gradient and temperature gradient, sigma A lambda_side dt/V_side . F] (sigma = -1 owner, +1
// other) for k_res_adjoint_gather.  A boundary face's viscous sweep goes through the ghost
// gradient map on the full gradients: its record's temperature-gradient adjoint is zero and the
// interior cell's density and pressure gradient adjoints go to t.gbx.  The face's area-vector
// sensitivity (inviscid) and a boundary face's inlet-state sensitivity are written here.
template <int DIM, bool VISC, bool BND>
__global__ __launch_bounds__(BLOCK, BND ? 1 : 2) void k_res_adjoint_face(AdArgs t)
{
    constexpr int NV = DIM + 2, NG = NV * DIM, NC = DIM * DIM + DIM, RS = NV + 1;
    const KArgs<double>& a = t.k;
    // two launches: two-sided faces [0, nfi), boundary faces [nfi, nf) (BND)
    const int f = (BND ? t.nfi : 0) + (int)(blockIdx.x * BLOCK + threadIdx.x);
    if (f >= (BND ? t.nf : t.nfi)) return;
    const int n = a.n;
    constexpr bool bnd = BND;
    const int cl = t.fcell[f], cr = bnd ? cl : t.fcell[(size_t)t.nf + f];
    const double g1d = a.ph.gamma - 1.0;
    double nv[DIM];
    face_n<DIM>(a.g, f, nv);
    const double A = a.g.fs[f];
    double qL[NV], qR[NV], lL[NV], lR[NV], mv[NV], qLb[NV], qRb[NV], gn[DIM], Fv[NV];
    load_q<DIM>(t, cl, qL);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        lL[v] = t.lrec[(size_t)cl * NV + v];
        lR[v] = bnd ? 0.0 : t.lrec[(size_t)cr * NV + v];
        mv[v] = A * (lR[v] - lL[v]);
        qLb[v] = 0.0;
        qRb[v] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < DIM; ++e) gn[e] = 0.0;
    BocoDev bc;
    if (bnd) {
        bc = a.boco[a.g.bboco[f - t.nfi]];
        bc_state<DIM>(bc, nv, qL, qR, a.ph);
    } else {
        load_q<DIM>(t, cr, qR);
    }
    hllc_rev<DIM>(a.ph.sch_hyp, a.ph.gamma, qL, qR, nv, mv, Fv, qLb, qRb, gn);
    // side records at the slot each side keeps the face in: [slot][cell][rs]; the volume share
    // (last entry) only when volume sensitivities are asked for
    const bool vol = t.gvol != nullptr;
    const int rs = vol ? RS : RS - 1;
    double* out = t.face + ((size_t)t.fslot[f] * n + cl) * rs;
    double* oto = bnd ? out : t.face + ((size_t)t.fslot[(size_t)t.nf + f] * n + cr) * rs;
    // gradients: full records on boundary faces (the ghost gradient map acts on them), viscous
    // records on two-sided faces
    constexpr int NB = VISC ? (BND ? NG : NC) : 1;
    double gLb[NB], gRb[NB];
    if (VISC) {
        double gL[NB], gR[NB], dk[DIM];
        const double* gsrc = BND ? a.grad : t.vgrad;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            gL[q] = gsrc[(size_t)cl * NB + q];
            gLb[q] = 0.0;
            gRb[q] = 0.0;
        }
        if (bnd) {
            bc_grad<DIM>(bc, nv, gL, gR);
        } else {
#pragma unroll
            for (int q = 0; q < NB; ++q) gR[q] = gsrc[(size_t)cr * NB + q];
        }
        const double dHL = t.fgeo[f];
#pragma unroll
        for (int e = 0; e < DIM; ++e) dk[e] = t.fgeo[(size_t)(1 + e) * t.nf + f];
        viscous_rev<DIM, !BND, !BND>(a.ph, qL, qR, dHL, dk, nv, gL, gR, mv, Fv, qLb, qRb, gLb, gRb);
        if (bnd) {
            // gLb += (dB_g/dgL)^T gRb: B_g copies the gradient, or reflects the scalar gradients
            // by S = I - 2 n n^T and the velocity gradient as S G S; both maps are self-adjoint
            double tb[NB];
            bc_grad<DIM>(bc, nv, gRb, tb);
#pragma unroll
            for (int q = 0; q < NB; ++q) gLb[q] += tb[q];
        }
    }
    if (vol) {
        double sL = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) sL += lL[v] * Fv[v];
        out[RS - 1] = -A * sL;
    }
    if (bnd) {
        const int bi = f - t.nfi;
        const bool inlet = bc.kind == BK_INLET_SUP;
        if (t.gq) {
#pragma unroll
            for (int v = 0; v < NV; ++v) t.gq[(size_t)v * t.nb + bi] = inlet ? qRb[v] : 0.0;
        }
        if (!inlet) {
            // qLb += (dB/dqL)^T qRb, gn += (dB/dn)^T qRb
            for (int e = 0; e < NV + DIM; ++e) {
                Dual qd[NV], nd[DIM], gd[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) qd[v] = Dual(qL[v], v == e ? 1.0 : 0.0);
#pragma unroll
                for (int d = 0; d < DIM; ++d) nd[d] = Dual(nv[d], NV + d == e ? 1.0 : 0.0);
                bc_state<DIM>(bc, nd, qd, gd, a.ph);
                double sacc = 0.0;
#pragma unroll
                for (int v = 0; v < NV; ++v) sacc += gd[v].d * qRb[v];
                if (e < NV)
                    qLb[e] += sacc;
                else
                    gn[e - NV] += sacc;
            }
        }
        if (VISC) {
#pragma unroll
            for (int q = 0; q < 2 * DIM; ++q) t.gbx[(size_t)bi * 2 * DIM + q] = gLb[q];
        }
    } else {
        double gRc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) gRc[v] = 0.0;
        prim_rev_q<DIM>(qR, g1d, qRb, gRc);
#pragma unroll
        for (int v = 0; v < NV; ++v) oto[v] = gRc[v];
        if (vol) {
            double sR = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) sR += lR[v] * Fv[v];
            oto[RS - 1] = A * sR;
        }
    }
    double gLc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) gLc[v] = 0.0;
    prim_rev_q<DIM>(qL, g1d, qLb, gLc);
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = gLc[v];
    if (VISC) {
        double* og = t.fgrad + ((size_t)t.fslot[f] * n + cl) * NC;
        if (BND) {
#pragma unroll
            for (int q = 0; q < DIM * DIM; ++q) og[q] = gLb[2 * DIM + q];
#pragma unroll
            for (int e = 0; e < DIM; ++e) og[DIM * DIM + e] = 0.0;
        } else {
#pragma unroll
            for (int q = 0; q < NC; ++q) og[q] = gLb[q];
        }
    }
    if (!VISC && t.gs) face_sens<DIM>(t, f, nv, A, mv, Fv, gn);
}

// per cell: its face sides' records summed in slot order (one writer per output, no atomics).
// Viscous: the summed temperature-gradient adjoint is chained through grad T = k1 grad p -
// k2 grad rho (k1 = 1/(rho r), k2 = p/(rho^2 r); linear in the adjoint, so once per cell), the
// boundary faces' density and pressure gradient adjoints are added, and the gradient adjoint gb of
// this cell is pushed back through g_c = sum_k w_k (q_k - q_c): this cell's own primitive adjoint
// gets -sum_k w_k . gb (and the ghost states' Jacobians on boundary slots, or the inlet-state
// sensitivity), and the far cell of slot k gets t_k = w_k . gb, left in t.slot for
// k_res_adjoint_tail.
template <int DIM, bool VISC, int NSM>
__global__ __launch_bounds__(BLOCK) void k_res_adjoint_gather(AdArgs t)
{
    constexpr int NV = DIM + 2, NG = NV * DIM, NC = DIM * DIM + DIM, RS = NV + 1;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    const bool vol = t.gvol != nullptr;
    const int rs = vol ? RS : RS - 1;
    double gu[NV], sv = 0.0, gc[VISC ? NC : 1], gx[VISC ? 2 * DIM : 1];
#pragma unroll
    for (int v = 0; v < NV; ++v) gu[v] = 0.0;
    if (VISC) {
#pragma unroll
        for (int q = 0; q < NC; ++q) gc[q] = 0.0;
#pragma unroll
        for (int q = 0; q < 2 * DIM; ++q) gx[q] = 0.0;
    }
    // slots unrolled to NSM >= ns: every slot's record loads can be in flight together
#pragma unroll
    for (int k = 0; k < NSM; ++k) {
        if (k >= a.ns) break;
        const int L = a.link[(size_t)k * n + c];
        if (L == SLOT_UNUSED) continue;
        const double* in = t.face + ((size_t)k * n + c) * rs;
#pragma unroll
        for (int v = 0; v < NV; ++v) gu[v] += in[v];
        if (VISC) {
            if (link_boundary(L)) {
                const double* fg = t.fgrad + ((size_t)k * n + c) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) gc[q] += fg[q];
                const double* bx = t.gbx + (size_t)link_index(L) * 2 * DIM;
#pragma unroll
                for (int q = 0; q < 2 * DIM; ++q) gx[q] += bx[q];
            } else {
                // the face's gradient adjoint, kept at its owner's slot, times this side's weight
                // (dHR = 1 - dHL on the owner's side, dHL on the other's)
                const bool oth = link_other(L);
                const double dHL = a.g.sh[(size_t)k * n + c];
                const double w = oth ? dHL : 1.0 - dHL;
                const double* fg = t.fgrad + (oth ? (size_t)link_mirror(L) * n + link_index(L) : (size_t)k * n + c) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) gc[q] += w * fg[q];
            }
        }
        if (vol) sv += in[RS - 1];
    }
    if (VISC) {
        const double g1d = a.ph.gamma - 1.0, r = a.ph.rgas;
        double qc[NV], qb[NV], gb[NG];
        load_q<DIM>(t, c, qc);
#pragma unroll
        for (int v = 0; v < NV; ++v) qb[v] = 0.0;
        // grad T = k1 grad p - k2 grad rho (calc_flux_viscous_cell.f90:96-109)
        const double k1 = 1.0 / (qc[0] * r), k2 = qc[1] / (qc[0] * qc[0] * r);
        const double* gcell = a.grad + (size_t)c * NG;
        double k1_b = 0.0, k2_b = 0.0;
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
            const double tb = gc[DIM * DIM + e];
            gb[e] = gx[e] - k2 * tb;
            gb[DIM + e] = gx[DIM + e] + k1 * tb;
            k1_b += tb * gcell[DIM + e];
            k2_b -= tb * gcell[e];
        }
#pragma unroll
        for (int q = 0; q < DIM * DIM; ++q) gb[2 * DIM + q] = gc[q];
        qb[1] += k2_b * k1 / qc[0];
        qb[0] += -k1_b * k1 / qc[0] - 2.0 * k2_b * k2 / qc[0];
        for (int k = 0; k < a.ns; ++k) {
            const int L = a.link[(size_t)k * n + c];
            if (L == SLOT_UNUSED) continue;
            double wk[DIM], tk[NV];
            slot_vec<DIM>(a.g.sw, k, n, c, wk);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double x = 0.0;
#pragma unroll
                for (int e = 0; e < DIM; ++e) x += wk[e] * gb[v * DIM + e];
                tk[v] = x;
                qb[v] -= x;
            }
            if (!link_boundary(L)) {
                double* o = t.slot + ((size_t)k * n + c) * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) o[v] = tk[v];
            } else {
                const int bi = link_index(L);
                if (a.boco[a.g.bboco[bi]].kind == BK_INLET_SUP) {
                    if (t.gq) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) t.gq[(size_t)v * t.nb + bi] += tk[v];
                    }
                } else {
                    // the ghost state's Jacobian is applied by k_res_adjoint_bnd_cells
                    double* o = t.slot + ((size_t)k * n + c) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) o[v] = tk[v];
                }
            }
        }
        prim_rev_q<DIM>(qc, g1d, qb, gu);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) t.gu[(size_t)v * n + c] = gu[v];
    // d(lambda.R_c)/dV_c = -lambda_c . R_c / V_c
    if (t.gvol) t.gvol[c] = -sv / a.vol[c];
}

// viscous, cells with boundary slots (t.bcell): the ghost cells' share of this cell's gradient
// adjoint, t.slot at the boundary slot, chained through the transpose of the ghost-state map's
// Jacobian (dual-number passes of the boundary template) and the primitive-variable reverse
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_res_adjoint_bnd_cells(AdArgs t)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int j0 = blockIdx.x * BLOCK + threadIdx.x;
    if (j0 >= t.nbcell) return;
    const int c = t.bcell[j0];
    const int n = a.n;
    const double g1d = a.ph.gamma - 1.0;
    double qc[NV], qb[NV];
    load_q<DIM>(t, c, qc);
#pragma unroll
    for (int v = 0; v < NV; ++v) qb[v] = 0.0;
    for (int k = 0; k < a.ns; ++k) {
        const int L = a.link[(size_t)k * n + c];
        if (L == SLOT_UNUSED || !link_boundary(L)) continue;
        const BocoDev bc = a.boco[a.g.bboco[link_index(L)]];
        if (bc.kind == BK_INLET_SUP) continue;
        const double* tk = t.slot + ((size_t)k * n + c) * NV;
        double nv[DIM];
        slot_normal<DIM>(a.g, k, n, c, nv);
        Dual nd[DIM];
#pragma unroll
        for (int e = 0; e < DIM; ++e) nd[e] = Dual(nv[e]);
        for (int j = 0; j < NV; ++j) {
            Dual qd[NV], gd[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) qd[v] = Dual(qc[v], v == j ? 1.0 : 0.0);
            bc_state<DIM>(bc, nd, qd, gd, a.ph);
            double sacc = 0.0;
#pragma unroll
            for (int v = 0; v < NV; ++v) sacc += gd[v].d * tk[v];
            qb[j] += sacc;
        }
    }
    double gu[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) gu[v] = 0.0;
    prim_rev_q<DIM>(qc, g1d, qb, gu);
#pragma unroll
    for (int v = 0; v < NV; ++v) t.gu[(size_t)v * n + c] += gu[v];
}

// viscous: each cell collects what its neighbours' gradients pushed back to it (t.slot at the
// mirror slot of the shared face), in slot order
template <int DIM>
__global__ __launch_bounds__(BLOCK) void k_res_adjoint_tail(AdArgs t)
{
    constexpr int NV = DIM + 2;
    const KArgs<double>& a = t.k;
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= a.n) return;
    const int n = a.n;
    double qb[NV], qc[NV], gu[NV];
    load_q<DIM>(t, c, qc);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        qb[v] = 0.0;
        gu[v] = 0.0;
    }
    for (int k = 0; k < a.ns; ++k) {
        const int L = a.link[(size_t)k * n + c];
        if (L == SLOT_UNUSED || link_boundary(L)) continue;
        const double* in = t.slot + ((size_t)link_mirror(L) * n + link_index(L)) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) qb[v] += in[v];
    }
    prim_rev_q<DIM>(qc, a.ph.gamma - 1.0, qb, gu);
#pragma unroll
    for (int v = 0; v < NV; ++v) t.gu[(size_t)v * n + c] += gu[v];
}

template <int DIM>
hipError_t launch_res_tangent(const AdArgs& t, hipStream_t s)
{
    const dim3 grid(grid_blocks(t.k.n));
    if (t.k.ph.viscous) {
        hipLaunchKernelGGL((k_grad_tangent<DIM>), grid, dim3(BLOCK), 0, s, t);
        hipLaunchKernelGGL((k_res_tangent_visc<DIM>), grid, dim3(BLOCK), 0, s, t);
    } else {
        hipLaunchKernelGGL((k_res_tangent<DIM>), grid, dim3(BLOCK), 0, s, t);
    }
    return hipGetLastError();
}

// t.fgeo from the slot geometry: once per scheme upload, before the first viscous adjoint launch
template <int DIM>
hipError_t launch_adj_face_geom(const AdArgs& t, hipStream_t s)
{
    hipLaunchKernelGGL((k_adj_face_geom<DIM>), dim3(grid_blocks(t.k.n)), dim3(BLOCK), 0, s, t);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_res_adjoint(const AdArgs& t, hipStream_t s)
{
    const bool rev = t.k.ph.sch_hyp == 25 || t.k.ph.sch_hyp == 26;
    const dim3 grid(grid_blocks(t.k.n)), fgrid((t.nfi + BLOCK - 1) / BLOCK), bgrid((t.nb + BLOCK - 1) / BLOCK);
    if (t.k.ph.viscous) {
        if (!rev) return hipErrorInvalidValue;  // the host rejects this case first
        if (t.nfi > 0) hipLaunchKernelGGL((k_res_adjoint_face<DIM, true, false>), fgrid, dim3(BLOCK), 0, s, t);
        if (t.nb > 0) hipLaunchKernelGGL((k_res_adjoint_face<DIM, true, true>), bgrid, dim3(BLOCK), 0, s, t);
        if (t.k.ns <= 4) hipLaunchKernelGGL((k_res_adjoint_gather<DIM, true, 4>), grid, dim3(BLOCK), 0, s, t);
        else if (t.k.ns <= 6) hipLaunchKernelGGL((k_res_adjoint_gather<DIM, true, 6>), grid, dim3(BLOCK), 0, s, t);
        else hipLaunchKernelGGL((k_res_adjoint_gather<DIM, true, MAX_SLOTS>), grid, dim3(BLOCK), 0, s, t);
        if (t.nbcell > 0)
            hipLaunchKernelGGL((k_res_adjoint_bnd_cells<DIM>), dim3((t.nbcell + BLOCK - 1) / BLOCK), dim3(BLOCK), 0,
                               s, t);
        hipLaunchKernelGGL((k_res_adjoint_tail<DIM>), grid, dim3(BLOCK), 0, s, t);
        return hipGetLastError();
    }
    if (!rev) {
        hipLaunchKernelGGL((k_res_adjoint<DIM>), grid, dim3(BLOCK), 0, s, t);
        return hipGetLastError();
    }
    if (t.nfi > 0) hipLaunchKernelGGL((k_res_adjoint_face<DIM, false, false>), fgrid, dim3(BLOCK), 0, s, t);
    if (t.nb > 0) hipLaunchKernelGGL((k_res_adjoint_face<DIM, false, true>), bgrid, dim3(BLOCK), 0, s, t);
    if (t.k.ns <= 4) hipLaunchKernelGGL((k_res_adjoint_gather<DIM, false, 4>), grid, dim3(BLOCK), 0, s, t);
    else if (t.k.ns <= 6) hipLaunchKernelGGL((k_res_adjoint_gather<DIM, false, 6>), grid, dim3(BLOCK), 0, s, t);
    else hipLaunchKernelGGL((k_res_adjoint_gather<DIM, false, MAX_SLOTS>), grid, dim3(BLOCK), 0, s, t);
    return hipGetLastError();
}

__global__ __launch_bounds__(BLOCK) void k_comb(double* out, Comb cb, long len)
{
    const long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= len) return;
    double s = 0.0;
    for (int i = 0; i < cb.m; ++i) s += cb.c[i] * cb.v[i][e];
    out[e] = s;
}

hipError_t launch_comb(double* out, const Comb& cb, long len, hipStream_t s)
{
    if (cb.m < 1 || cb.m > COMB_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_comb, dim3((unsigned)((len + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, out, cb, len);
    return hipGetLastError();
}

// the same sums over [nv][n] vectors, written as one record per cell scaled by dt / V_c (the
// lambda dt / V records of the HLLC adjoint kernels)
__global__ __launch_bounds__(BLOCK) void k_comb_lrec(double* out, Comb cb, const double* vol, const double* d_dt,
                                                     int nv, int n)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= n) return;
    const double k = *d_dt / vol[c];
    double s[NVMAX];
#pragma unroll
    for (int v = 0; v < NVMAX; ++v) s[v] = 0.0;
    for (int i = 0; i < cb.m; ++i)
#pragma unroll
        for (int v = 0; v < NVMAX; ++v)
            if (v < nv) s[v] += cb.c[i] * cb.v[i][(size_t)v * n + c];
#pragma unroll
    for (int v = 0; v < NVMAX; ++v)
        if (v < nv) out[(size_t)c * nv + v] = s[v] * k;
}

hipError_t launch_comb_lrec(double* lrec, const Comb& cb, const double* vol, const double* d_dt, int nv, int n,
                            hipStream_t s)
{
    if (cb.m < 1 || cb.m > COMB_MAX || nv > NVMAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_comb_lrec, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, lrec, cb, vol, d_dt,
                       nv, n);
    return hipGetLastError();
}

template hipError_t launch_res_tangent<2>(const AdArgs&, hipStream_t);
template hipError_t launch_res_tangent<3>(const AdArgs&, hipStream_t);
template hipError_t launch_adj_face_geom<2>(const AdArgs&, hipStream_t);
template hipError_t launch_adj_face_geom<3>(const AdArgs&, hipStream_t);
template hipError_t launch_res_adjoint<2>(const AdArgs&, hipStream_t);
template hipError_t launch_res_adjoint<3>(const AdArgs&, hipStream_t);

}  // namespace ty

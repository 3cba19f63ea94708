// Reverse-mode (adjoint) differentiation of the HLLC / HLLCB flux, written by hand.
//
// hllc_rev evaluates the flux F(qL, qR, n) exactly as num_flux does for sch 25/26
// (calc_flux_hllc.f90:61-171 with calc_roe_states.f90:35-53) and, given a weight
// Fb = dJ/dF, accumulates qLb += (dF/dqL)^T Fb, qRb += (dF/dqR)^T Fb and
// nb += (dF/dn)^T Fb in one forward and one reverse sweep: the cost of a few fluxes
// instead of one dual-number pass per input direction.  Branches (upwind sides,
// the side of S*, the min/max of the wave-speed estimates) follow the primal.
#pragma once
#include "ty_kernels_impl.h"

namespace ty {

template <int DIM>
__device__ __forceinline__ void hllc_rev(int sch, double g, const double* qL, const double* qR, const double* n,
                                         const double* Fb, double* F, double* qLb, double* qRb, double* nb)
{
    constexpr int NV = DIM + 2;
    const double ig1 = 1.0 / (g - 1.0);
    const double gig1 = g / (g - 1.0);
    const double beta = (sch == 26) ? ::sqrt((g - 1.0) / (2.0 * g)) : 1.0;
    const double* uL = qL + 2;
    const double* uR = qR + 2;
    // ---- forward (same operations as num_flux / roe_avg)
    const double vnl = dotd<DIM>(uL, n), vnr = dotd<DIM>(uR, n);
    const double uL2 = dotd<DIM>(uL, uL), uR2 = dotd<DIM>(uR, uR);
    const double rel = ig1 * qL[1] + 0.5 * qL[0] * uL2;
    const double rer = ig1 * qR[1] + 0.5 * qR[0] * uR2;
    const double irL = 1.0 / qL[0], irR = 1.0 / qR[0];
    const double t1 = ::sqrt(qR[0] * irL);
    const double kl = 1.0 / (1.0 + t1);
    const double kr = t1 * kl;
    double rv[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) rv[d] = kl * uL[d] + kr * uR[d];
    const double hl = gig1 * qL[1] * irL + 0.5 * uL2;
    const double hr = gig1 * qR[1] * irR + 0.5 * uR2;
    const double h = kl * hl + kr * hr;
    const double rv2 = dotd<DIM>(rv, rv);
    const double a2 = (g - 1.0) * (h - 0.5 * rv2);
    const double vm = dotd<DIM>(rv, n);
    const double am = ::sqrt(a2);
    const double al = ::sqrt(g * qL[1] * irL), ar = ::sqrt(g * qR[1] * irR);
    const double slA = vnl - beta * al, slB = vm - am;
    const double srA = vnr + beta * ar, srB = vm + am;
    const bool slFirst = slA < slB, srFirst = srA > srB;
    const double sl = slFirst ? slA : slB;
    const double sr = srFirst ? srA : srB;
    // adjoints of the intermediates
    double vnlb = 0, vnrb = 0, relb = 0, rerb = 0, slb = 0, srb = 0;
    if (sl >= 0.0 || sr <= 0.0) {
        // upwind: F = F(q_side)
        const bool Lside = sl >= 0.0;
        const double* q = Lside ? qL : qR;
        const double* u = q + 2;
        const double vn = Lside ? vnl : vnr;
        const double re = Lside ? rel : rer;
        double* qb = Lside ? qLb : qRb;
        F[0] = vn * q[0];
        F[1] = vn * (re + q[1]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[2 + d] = (vn * q[0]) * u[d] + q[1] * n[d];
        double vnb = Fb[0] * q[0] + Fb[1] * (re + q[1]);
        double reb = Fb[1] * vn;
        qb[0] += Fb[0] * vn;
        qb[1] += Fb[1] * vn;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            vnb += Fb[2 + d] * q[0] * u[d];
            qb[0] += Fb[2 + d] * vn * u[d];
            qb[2 + d] += Fb[2 + d] * vn * q[0];
            qb[1] += Fb[2 + d] * n[d];
            nb[d] += Fb[2 + d] * q[1];
        }
        if (Lside) { vnlb += vnb; relb += reb; } else { vnrb += vnb; rerb += reb; }
    } else {
        const double rqL = qL[0] * (sl - vnl), rqR = qR[0] * (sr - vnr);
        const double num = rqR * vnr - rqL * vnl - qR[1] + qL[1];
        const double den = rqR - rqL;
        const double sst = num / den;
        const bool left = sst >= 0.0;
        const double* q = left ? qL : qR;
        const double* u = q + 2;
        const double rq = left ? rqL : rqR;
        const double vn = left ? vnl : vnr;
        const double s = left ? sl : sr;
        const double re = left ? rel : rer;
        double* qb = left ? qLb : qRb;
        const double iks = 1.0 / (s - sst);
        const double rst = rq * iks;
        const double pst = q[1] - rq * (vn - sst);
        const double t2 = re * (s - vn) - q[1] * vn + pst * sst;
        const double rest = iks * t2;
        double tmp[DIM], rvst[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            tmp[d] = (rq * u[d]) + (pst - q[1]) * n[d];
            rvst[d] = iks * tmp[d];
        }
        F[0] = sst * rst;
        F[1] = sst * (rest + pst);
#pragma unroll
        for (int d = 0; d < DIM; ++d) F[2 + d] = (sst * rvst[d]) + (pst * n[d]);
        // reverse
        double sstb = Fb[0] * rst + Fb[1] * (rest + pst);
        const double rstb = Fb[0] * sst;
        const double restb = Fb[1] * sst;
        double pstb = Fb[1] * sst;
        double iksb = 0, rqb = 0, vnb = 0, sb = 0, reb = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            sstb += Fb[2 + d] * rvst[d];
            pstb += Fb[2 + d] * n[d];
            nb[d] += Fb[2 + d] * pst;
            const double rvb = Fb[2 + d] * sst;
            iksb += rvb * tmp[d];
            const double tb = rvb * iks;
            rqb += tb * u[d];
            qb[2 + d] += tb * rq;
            pstb += tb * n[d];
            qb[1] -= tb * n[d];
            nb[d] += tb * (pst - q[1]);
        }
        // rest = iks * t2, t2 = re (s - vn) - p vn + pst sst
        iksb += restb * t2;
        const double t2b = restb * iks;
        reb += t2b * (s - vn);
        sb += t2b * re;
        vnb -= t2b * re;
        qb[1] -= t2b * vn;
        vnb -= t2b * q[1];
        pstb += t2b * sst;
        sstb += t2b * pst;
        // pst = p - rq (vn - sst)
        qb[1] += pstb;
        rqb -= pstb * (vn - sst);
        vnb -= pstb * rq;
        sstb += pstb * rq;
        // rst = rq iks
        rqb += rstb * iks;
        iksb += rstb * rq;
        // iks = 1 / (s - sst)
        sb -= iksb * iks * iks;
        sstb += iksb * iks * iks;
        // sst = num / den
        const double numb = sstb / den;
        const double denb = -sstb * sst / den;
        double rqLb = -numb * vnl - denb;
        double rqRb = numb * vnr + denb;
        vnrb += numb * rqR;
        vnlb -= numb * rqL;
        qRb[1] -= numb;
        qLb[1] += numb;
        if (left) { rqLb += rqb; vnlb += vnb; slb += sb; relb += reb; }
        else { rqRb += rqb; vnrb += vnb; srb += sb; rerb += reb; }
        // rqL = rhoL (sl - vnl), rqR = rhoR (sr - vnr)
        qLb[0] += rqLb * (sl - vnl);
        slb += rqLb * qL[0];
        vnlb -= rqLb * qL[0];
        qRb[0] += rqRb * (sr - vnr);
        srb += rqRb * qR[0];
        vnrb -= rqRb * qR[0];
    }
    // wave speeds
    double alb = 0, arb = 0, vmb = 0, amb = 0;
    if (slFirst) { vnlb += slb; alb -= beta * slb; } else { vmb += slb; amb -= slb; }
    if (srFirst) { vnrb += srb; arb += beta * srb; } else { vmb += srb; amb += srb; }
    // al = sqrt(g pL / rhoL), ar likewise
    if (alb != 0.0) {
        const double c = alb * g / (2.0 * al);
        qLb[1] += c * irL;
        qLb[0] -= c * qL[1] * irL * irL;
    }
    if (arb != 0.0) {
        const double c = arb * g / (2.0 * ar);
        qRb[1] += c * irR;
        qRb[0] -= c * qR[1] * irR * irR;
    }
    // am = sqrt(a2), a2 = (g - 1)(h - |rv|^2 / 2), vm = rv . n
    double hb = 0.0, rv2b = 0.0;
    if (amb != 0.0) {
        const double a2b = amb / (2.0 * am);
        hb += (g - 1.0) * a2b;
        rv2b -= 0.5 * (g - 1.0) * a2b;
    }
    double rvb[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        rvb[d] = vmb * n[d] + 2.0 * rv[d] * rv2b;
        nb[d] += vmb * rv[d];
    }
    // h = kl hl + kr hr, hl = gig1 pL / rhoL + |uL|^2 / 2, hr likewise
    double klb = hb * hl, krb = hb * hr;
    const double hlb = hb * kl, hrb = hb * kr;
    qLb[1] += hlb * gig1 * irL;
    qLb[0] -= hlb * gig1 * qL[1] * irL * irL;
    qRb[1] += hrb * gig1 * irR;
    qRb[0] -= hrb * gig1 * qR[1] * irR * irR;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        qLb[2 + d] += hlb * uL[d];
        qRb[2 + d] += hrb * uR[d];
        // rv = kl uL + kr uR
        klb += rvb[d] * uL[d];
        krb += rvb[d] * uR[d];
        qLb[2 + d] += rvb[d] * kl;
        qRb[2 + d] += rvb[d] * kr;
    }
    // kr = t1 kl, kl = 1 / (1 + t1), t1 = sqrt(rhoR / rhoL)
    double t1b = krb * kl;
    klb += krb * t1;
    t1b -= klb * kl * kl;
    const double c1 = t1b / (2.0 * t1);
    qRb[0] += c1 * irL;
    qLb[0] -= c1 * qR[0] * irL * irL;
    // rel = ig1 pL + rhoL |uL|^2 / 2, rer likewise; vnl = uL . n, vnr likewise
    qLb[1] += relb * ig1;
    qLb[0] += relb * 0.5 * uL2;
    qRb[1] += rerb * ig1;
    qRb[0] += rerb * 0.5 * uR2;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        qLb[2 + d] += relb * qL[0] * uL[d] + vnlb * n[d];
        qRb[2 + d] += rerb * qR[0] * uR[d] + vnrb * n[d];
        nb[d] += vnlb * uL[d] + vnrb * uR[d];
    }
    (void)NV;
}

// Adjoint of calc_varprim_ns (calc_varprim_ns.f90:47-52): Ub += (dq/dU)^T qb for
// q = (rho, p, u) of U = (rho, rhoE, rho u).
template <int DIM>
__device__ __forceinline__ void prim_rev(const double* U, double g1, const double* qb, double* Ub)
{
    const double rho = U[0];
    const double ir = 1.0 / rho;
    double u[DIM], uu = 0.0, ub_u = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        u[d] = U[2 + d] * ir;
        uu += u[d] * u[d];
        ub_u += qb[2 + d] * u[d];
    }
    Ub[0] += qb[0] + qb[1] * g1 * 0.5 * uu - ub_u * ir;
    Ub[1] += qb[1] * g1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) Ub[2 + d] += -g1 * u[d] * qb[1] + qb[2 + d] * ir;
}

}  // namespace ty

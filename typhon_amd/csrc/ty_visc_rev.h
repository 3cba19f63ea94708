// Reverse sweep of the cell-average / cell-full viscous face flux (viscous_face in
// ty_kernels_impl.h; calc_flux_viscous_cell.f90:75-158 with the Sutherland law of
// EQNS.f90:225-265).  Given the weight Fb = dJ/dF it adds to F the viscous part of the
// flux (the forward value, same operations as viscous_face) and accumulates
// qLb += (dF/dqL)^T Fb, qRb += (dF/dqR)^T Fb, gLb += (dF/dgL)^T Fb, gRb += (dF/dgR)^T Fb.
// The face geometry (distance weight dHL, dHR = 1 - dHL, vLR, n) is held fixed: this is the
// state adjoint.  Gradient records: [rho (DIM), p (DIM), d u_i / d x_j at (2 + i) DIM + j].
// COMPACT: each side's gradients come as its viscous record [d u_i / d x_j (DIM^2), grad T (DIM)]
// (grad T = k1 grad p - k2 grad rho, k1 = 1/(rho r), k2 = p/(rho^2 r), formed by the gradient
// kernel) and gLb / gRb are the adjoints of those records; qLb / qRb then leave out grad T's
// dependence on (rho, p), which the per-cell gather adds once per cell from the summed grad T
// adjoint (it is linear in it).
#pragma once
#include "ty_kernels_impl.h"

namespace ty {

// SHARED (with COMPACT): gLb receives the adjoint of the face's averaged velocity gradient G0 and
// temperature gradient (assigned, not weighted) and gRb is left alone: each side's adjoint is its
// distance weight (dHR for L, dHL for R) times it.
template <int DIM, bool COMPACT = false, bool SHARED = false>
__device__ __forceinline__ void viscous_rev(const Physics& ph, const double* qL, const double* qR, double dHL,
                                            const double* vLR, const double* nf, const double* gL,
                                            const double* gR, const double* Fb, double* F, double* qLb, double* qRb,
                                            double* gLb, double* gRb)
{
    // ---- forward
    const double r = ph.rgas;
    const double cp = ph.gamma * r / (ph.gamma - 1.0);
    const double dHR = 1.0 - dHL;
    double velH[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) velH[e] = dHR * qL[2 + e] + dHL * qR[2 + e];
    const double TL = qL[1] / (qL[0] * r), TR = qR[1] / (qR[0] * r);
    const double TH = dHR * TL + dHL * TR;
    const double rhoH = dHR * qL[0] + dHL * qR[0];
    const double kl1 = 1.0 / (qL[0] * r), kl2 = qL[1] / (qL[0] * qL[0] * r);
    const double kr1 = 1.0 / (qR[0] * r), kr2 = qR[1] / (qR[0] * qR[0] * r);
    constexpr int GU = COMPACT ? 0 : 2 * DIM;  // velocity-gradient offset in the gradient records
    double gTL[DIM], gTR[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        gTL[e] = COMPACT ? gL[DIM * DIM + e] : kl1 * gL[DIM + e] - kl2 * gL[e];
        gTR[e] = COMPACT ? gR[DIM * DIM + e] : kr1 * gR[DIM + e] - kr2 * gR[e];
    }
    double mu, dmu_dTH = 0.0, dmu_drhoH = 0.0;
    if (ph.visc_law == 11) {
        mu = ph.visc_dyn * dpow(TH / ph.tref, 1.5) * (ph.tref + ph.tsuth) / (TH + ph.tsuth);
        dmu_dTH = mu * (1.5 / TH - 1.0 / (TH + ph.tsuth));
    } else if (ph.visc_law == 14) {
        mu = ph.visc_kin * rhoH;
        dmu_drhoH = ph.visc_kin;
    } else {
        mu = ph.visc_dyn;
    }
    double G0[DIM][DIM], G[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) G0[i][j] = dHL * gR[GU + i * DIM + j] + dHR * gL[GU + i * DIM + j];
    const double dLR2 = dotd<DIM>(vLR, vLR);
    double gavg[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) gavg[e] = dHL * gTR[e] + dHR * gTL[e];
    const bool full = ph.sch_dis == 10;
    double kv[DIM], vi[DIM], pscal = 0.0, gTH;
#pragma unroll
    for (int e = 0; e < DIM; ++e) kv[e] = (1.0 / dLR2) * vLR[e];
    if (full) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const double fa = qR[2 + i] - qL[2 + i] - dotd<DIM>(G0[i], vLR);
#pragma unroll
            for (int j = 0; j < DIM; ++j) G[i][j] = G0[i][j] + fa * kv[j];
        }
        pscal = dotd<DIM>(vLR, nf) / dLR2;
        const double fcomp = pscal * (TR - TL);
#pragma unroll
        for (int e = 0; e < DIM; ++e) vi[e] = nf[e] - pscal * vLR[e];
        gTH = fcomp + dotd<DIM>(gavg, vi);
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) G[i][j] = G0[i][j];
#pragma unroll
        for (int e = 0; e < DIM; ++e) vi[e] = nf[e];
        gTH = dotd<DIM>(gavg, nf);
    }
    double tr = G[0][0];
#pragma unroll
    for (int i = 1; i < DIM; ++i) tr += G[i][i];
    const double addv = -(2.0 / 3.0) * tr;
    double sn[DIM], s1[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double s = 0.0, su = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            double sij = G[i][j] + G[j][i];
            if (i == j) sij += addv;
            s += (mu * sij) * nf[j];
            su += sij * nf[j];
        }
        sn[i] = s;
        s1[i] = su;
    }
    const double svn = dotd<DIM>(sn, velH);
    const double conduct = mu * cp / ph.prandtl;
#pragma unroll
    for (int e = 0; e < DIM; ++e) F[2 + e] = F[2 + e] - sn[e];
    F[1] = F[1] - svn;
    F[1] = F[1] - conduct * gTH;

    // ---- reverse
    const double svn_b = -Fb[1];
    const double conduct_b = -Fb[1] * gTH;
    const double gTH_b = -Fb[1] * conduct;
    double sn_b[DIM], velH_b[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        sn_b[e] = -Fb[2 + e] + svn_b * velH[e];
        velH_b[e] = svn_b * sn[e];
    }
    double mu_b = conduct_b * cp / ph.prandtl + dotd<DIM>(sn_b, s1);
    // sn_i = mu sum_j (G_ij + G_ji + delta_ij addv) n_j
    double Gb[DIM][DIM];
    double addv_b = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) addv_b += mu * sn_b[i] * nf[i];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) Gb[i][j] = mu * (sn_b[i] * nf[j] + sn_b[j] * nf[i]);
    const double tr_b = -(2.0 / 3.0) * addv_b;
#pragma unroll
    for (int i = 0; i < DIM; ++i) Gb[i][i] += tr_b;
    double gavg_b[DIM], TL_b = 0.0, TR_b = 0.0;
#pragma unroll
    for (int e = 0; e < DIM; ++e) gavg_b[e] = gTH_b * vi[e];
    double uLb[DIM], uRb[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        uLb[e] = 0.0;
        uRb[e] = 0.0;
    }
    if (full) {
        TR_b += pscal * gTH_b;
        TL_b -= pscal * gTH_b;
        // G = G0 + fa kv^T, fa_i = uR_i - uL_i - G0_i . vLR
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const double fa_b = dotd<DIM>(Gb[i], kv);
            uRb[i] += fa_b;
            uLb[i] -= fa_b;
#pragma unroll
            for (int j = 0; j < DIM; ++j) Gb[i][j] -= fa_b * vLR[j];
        }
    }
    // G0 = dHL gR_u + dHR gL_u
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            if (SHARED) {
                gLb[GU + i * DIM + j] = Gb[i][j];
            } else {
                gRb[GU + i * DIM + j] += dHL * Gb[i][j];
                gLb[GU + i * DIM + j] += dHR * Gb[i][j];
            }
        }
    // gavg = dHL gTR + dHR gTL; gT = k1 grad p - k2 grad rho
    double kl1_b = 0.0, kl2_b = 0.0, kr1_b = 0.0, kr2_b = 0.0;
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        const double tLb = dHR * gavg_b[e], tRb = dHL * gavg_b[e];
        if (SHARED) {
            gLb[DIM * DIM + e] = gavg_b[e];
        } else if (COMPACT) {
            gLb[DIM * DIM + e] += tLb;
            gRb[DIM * DIM + e] += tRb;
        } else {
            gLb[DIM + e] += kl1 * tLb;
            gLb[e] -= kl2 * tLb;
            gRb[DIM + e] += kr1 * tRb;
            gRb[e] -= kr2 * tRb;
            kl1_b += tLb * gL[DIM + e];
            kl2_b -= tLb * gL[e];
            kr1_b += tRb * gR[DIM + e];
            kr2_b -= tRb * gR[e];
        }
    }
    const double TH_b = mu_b * dmu_dTH;
    const double rhoH_b = mu_b * dmu_drhoH;
    TL_b += dHR * TH_b;
    TR_b += dHL * TH_b;
    // kl1 = 1/(rho r), kl2 = p/(rho^2 r), T = p/(rho r) (kl1_b = kl2_b = 0 when COMPACT)
    {
        const double iL = 1.0 / (qL[0] * r), iR = 1.0 / (qR[0] * r);
        qLb[1] += TL_b * iL + kl2_b * iL / qL[0];
        qLb[0] += -TL_b * TL / qL[0] - kl1_b * kl1 / qL[0] - 2.0 * kl2_b * kl2 / qL[0] + dHR * rhoH_b;
        qRb[1] += TR_b * iR + kr2_b * iR / qR[0];
        qRb[0] += -TR_b * TR / qR[0] - kr1_b * kr1 / qR[0] - 2.0 * kr2_b * kr2 / qR[0] + dHL * rhoH_b;
    }
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        qLb[2 + e] += dHR * velH_b[e] + uLb[e];
        qRb[2 + e] += dHL * velH_b[e] + uRb[e];
    }
}

}  // namespace ty

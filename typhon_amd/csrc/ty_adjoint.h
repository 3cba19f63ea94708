// Arguments and launchers of the tangent-linear / adjoint residual kernels
// (ty_adjoint.hip).  Device arrays, device cell and face order.
#pragma once
#include "ty_internal.h"

namespace ty {

struct AdArgs {
    KArgs<double> k;      // mesh, geometry, boundary conditions, physics; k.us = FP64 state [nv][n]
    int nf, nb;           // device faces, device boundary faces
    const int* fcell;     // [2][nf] owner and other cell of each device face (other = -1 on boundary faces)
    int nfi;              // two-sided faces come first: device faces [0, nfi)
    double* face;         // [ns][n][RS] work: per face side, at the slot where that side's cell keeps
                          // the face, the cell's state adjoint and (with volume sensitivities) the
                          // side's share of its volume sensitivity; RS = nv (+ 1)
    double* fgrad;        // [ns][n][dim^2 + dim] work (viscous): at the owner's slot of a two-sided
                          // face, the adjoint of the face's averaged velocity gradient and temperature
                          // gradient (each side's gradient adjoint is its distance weight times it); at
                          // a boundary face's slot, the interior cell's velocity-gradient adjoint
    const unsigned char* fslot;  // [2][nf] slot of each device face in its owner's / other cell's list
    const double* qrec;   // [n][nv] primitive records (rho, p, u..) of k.us (the HLLC face kernels)
    double* lrec;         // [n][nv] lambda dt / V per cell, one record per cell (HLLC face kernels),
                          // filled by the caller
    // viscous residuals (k.grad holds the cell gradients of the state, vgrad their viscous records
    // [n][dim^2 + dim]: velocity gradient and temperature gradient)
    const double* vgrad;
    const double* fgeo;   // [4][nf] work: the viscous distance weight dHL and the canonical centre
                          // difference d per face
    double* gbx;          // [nb][2 dim] work: a boundary face's adjoint of its interior cell's density
                          // and pressure gradients (the ghost side's gradient map included)
    double* slot;         // [ns][n][nv] work: per slot, the far cell's primitive adjoint through this
                          // cell's gradient
    double* dgrad;        // [n][nv dim] work: tangent of the cell gradients
    // tangent inputs (each may be null = zero direction) and output
    const double* du;     // [nv][n] state direction
    const double* ds;     // [3][nf] area-vector direction per device face (owner orientation)
    const double* dvol;   // [n] volume direction
    const double* dq;     // [nboco][NVMAX] supersonic-inlet ghost primitive state direction
    double* out;          // [nv][n] dR
    // adjoint input and outputs (gs, gvol, gq may be null)
    const double* lam;    // [nv][n]
    double* gu;           // [nv][n] (dR/dU)^T lambda
    double* gs;           // [3][nf] (dR/dS)^T lambda, written by each face's owner cell
    double* gvol;         // [n] (dR/dV)^T lambda
    double* gq;           // [nv][nb] per boundary face: (dR/dq_ghost)^T lambda (inlet faces, else 0)
    const int* bcell;     // [nbcell] cells with a boundary slot, ascending
    int nbcell;
};

// out = sum_i c[i] v[i] over m vectors of len doubles, summed in index order (RK stage states and
// the reverse RK sweep); out may alias one of the v[i]
constexpr int COMB_MAX = MAX_STAGES + 1;
struct Comb {
    const double* v[COMB_MAX];
    double c[COMB_MAX];
    int m;
};
hipError_t launch_comb(double* out, const Comb& cb, long len, hipStream_t s);
// lrec[c][v] = (sum_i c[i] v[i][v][c]) dt / vol[c]: the combination as lambda dt / V records
hipError_t launch_comb_lrec(double* lrec, const Comb& cb, const double* vol, const double* d_dt, int nv, int n,
                            hipStream_t s);

template <int DIM>
hipError_t launch_res_tangent(const AdArgs& t, hipStream_t s);
template <int DIM>
hipError_t launch_res_adjoint(const AdArgs& t, hipStream_t s);
template <int DIM>
hipError_t launch_adj_face_geom(const AdArgs& t, hipStream_t s);  // fills t.fgeo (viscous adjoint)

}  // namespace ty

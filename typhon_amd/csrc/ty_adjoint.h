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
    double* face;         // [nf][2][nv + 1] work: per face side, that cell's state adjoint and the
                          // side's share of its volume sensitivity
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
};

template <int DIM>
hipError_t launch_res_tangent(const AdArgs& t, hipStream_t s);
template <int DIM>
hipError_t launch_res_adjoint(const AdArgs& t, hipStream_t s);

}  // namespace ty

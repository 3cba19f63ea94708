// Internal declarations shared by the HIP kernels (ty_kernels_*.hip) and the
// host side of the C-ABI (ty_host.cpp).  gfx950 only.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ty {

constexpr int NVMAX = 5;         // conserved variables (rho, rhoE, rho u, rho v, rho w)
constexpr int MAX_SLOTS = 8;     // faces per cell
constexpr int MAX_STAGES = 8;    // RK stages
constexpr int FACE_BITS = 26;    // face id bits of a slot word
constexpr uint32_t FACE_MASK = (1u << FACE_BITS) - 1u;
constexpr uint32_t PERIODIC_BIT = 1u << FACE_BITS;  // face joins periodic partners
constexpr int MIRROR_SHIFT = 27;  // bits 27-29: slot of the same face in the far cell's list
constexpr int SLOT_UNUSED = -1;  // slot word of an empty slot
constexpr int BLOCK = 256;       // threads per block (4 wavefronts of 64)
constexpr int MAX_BOCO = 64;

// Reciprocal, quotient and square root in the arithmetic precision of the residual kernels:
// IEEE-rounded in FP64 (the reference's arithmetic); in the FP32 form of the mixed mode the
// hardware v_rcp_f32 / v_sqrt_f32 (about 1 ulp) instead of the correctly rounded sequences.
template <typename R>
__device__ __forceinline__ R arcp(R x) { return R(1) / x; }
template <typename R>
__device__ __forceinline__ R adiv(R a, R b) { return a / b; }
template <typename R>
__device__ __forceinline__ R asqrt(R x) { return sqrt(x); }
template <>
__device__ __forceinline__ float arcp(float x) { return __builtin_amdgcn_rcpf(x); }
template <>
__device__ __forceinline__ float adiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <>
__device__ __forceinline__ float asqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// boundary kinds handled on the device (the reference's bc_* codes map onto these)
enum BcKind : int {
    BK_SYM = 0, BK_INLET_SUP = 1, BK_OUTLET_SUB = 2, BK_OUTLET_SUP = 3,
    BK_WALL_ISOTH = 4, BK_WALL_FLUX = 5, BK_INLET_SUB = 6
};

struct BocoDev {
    int kind;
    int pad;
    double param[8];  // BK_INLET_SUP: precomputed ghost primitive state (rho, p, u, v, w);
                      // BK_OUTLET_SUB: param[0] = imposed static pressure;
                      // BK_WALL_ISOTH: param[0] = wall temperature, param[4..6] = wall velocity;
                      // BK_WALL_FLUX: param[0] = wall heat flux (outward), param[4..6] = wall velocity;
                      // BK_INLET_SUB: param[0] = total pressure, param[1] = total temperature,
                      //               param[3..5] = unit flow direction
};

// Device geometry in the arithmetic precision R.  The residual kernels read it
// per face slot, [slot][component][cell], so every load of a wavefront is
// coalesced and nothing geometric is gathered; the tangent/adjoint kernels read
// the per-face arrays.  "Canonical" quantities are those of the face's owner
// (the reference's ic1 side): the normal points owner -> other and
// d = x_R - x_L is taken in the owner's frame (periodic far sides translated).
// four values of precision R per record: FP32 as one float4 array, FP64 as two planes of double2
// (x, y | z, w), so that each load of a wavefront reads one contiguous block
template <typename R>
struct Pack4;
template <>
struct Pack4<float> {
    using S = float4;
    static constexpr int planes = 1;
};
template <>
struct Pack4<double> {
    using S = double2;
    static constexpr int planes = 2;
};

template <typename R>
struct Geo {
    const R* ss;    // [ns][3][n] canonical area vector S = A n of the slot's face
    const R* sw;    // [ns][3][n] LSQ weights: grad q_c = sum_k w_k (q_k - q_c)
    const R* sr;    // [ns][3][n] face centre - cell centre, in this cell's frame
    const R* sd;    // [ns][3][n] canonical x_R - x_L (the ghost centre on boundary faces)
    const R *fnx, *fny, *fnz, *fs;  // [nf] per device face: normal, area
    const int* bboco;               // [nb] boco index of each boundary face
    const R* eps2;                  // [n] Venkatakrishnan (K h)^3
    const R* sh;                    // [ns][n] viscous face weight dHL = |x_f - x_L| / (|x_f - x_L| + |x_f - x_R|)
                                    // (canonical sides; the other side's weight is 1 - dHL)
    // the same data packed per slot, [ns][n] records of four R (Pack4), read in the interior paths
    // of the FP32 viscous flux (pa = (S, dHL), pb = (x_R - x_L, link word); null in FP64) and of
    // the plain-gradient kernels (pw = (LSQ weight, link word)); the link word's bits are stored in
    // the fourth element
    const typename Pack4<R>::S* pa;
    const typename Pack4<R>::S* pb;
    const typename Pack4<R>::S* pw;
};

struct Physics {
    double gamma, rgas, prandtl;
    int visc_law;
    double visc_dyn, visc_kin, tref, tsuth;
    int viscous;
    int sch_dis;
    int sch_hyp;
    char method;   // 'N','M','F','U'
    char limiter;  // 'X','M','A','V','S','K','L','W'
    int lsqw;      // weighted LSQ
    int barth;     // Barth / super-Barth post-limiter on
    double klim;
    int mono;      // monotonic post-limiter 10, 11, 12 (MONOTONIC0/1/2) of second-order face states; 0 off
    // launch constants derived on the host: 1/(gamma-1), gamma/(gamma-1), 1/r,
    // conductivity factor cp/Pr, 1/Tref and Tref + S of Sutherland's law
    double ig1, gig1, irgas, cp_pr, itref, tsum;
};

// Everything one launch of the residual kernels needs, passed by value.
template <typename R>
struct KArgs {
    int n, ns, nblk_real;
    int dense;       // every cell uses all ns slots
    const int* nbr;  // [ns][n] >=0 neighbour cell, <0: -(boundary face + 1)
    const int* fw;   // [ns][n] slot word: face id | periodic << 26 | mirror slot << 27 | other-side << 31;
                     // -1 unused (tangent/adjoint kernels)
    const int* link; // [ns][n] link word of the residual kernels: boundary << 31 | other-side << 30 |
                     // mirror slot << 27 | neighbour cell or boundary face; -1 unused
    const unsigned* bnd_bits;  // bit g of word g / 32: cells [32 g, 32 g + 32) include one with a boundary
                               // slot (null: assume every group does)
    Geo<R> g;
    const BocoDev* boco;
    const double* vol;    // [n]
    Physics ph;
    // state
    const R* us;          // [n][nv] stage state as primitive records (rho, p, u..), one per cell,
                          // recovered from the FP64 conservative state by its producer
    R* us_next;           // [n][nv]
    double* u0;           // [nv][n]
    double* acc;          // [nv][n] the step's sum of b_j dt/V R_j over the stages so far
    float* acc32;         // non-null (FP32 kernels of TY_PREC_MIXED_F32SUM): that sum kept in FP32 here
    double* rstore;       // [nstore][nv][n]
    const R* grad;        // [n][nv*dim] LSQ gradients of (rho, p, u..), one record per cell; with vrec
                          // set, the viscous records [n][dim*dim + dim] instead: the velocity
                          // gradient (row-major d u_i / d x_j) and the temperature gradient
    R* grad_out;          // written by K1 when not null
    int vrec;             // K1 writes (and K2 reads) viscous records
    R* vrec_out;          // with vrec unset: K1 also writes each cell's viscous record here (adjoint)
    const R* gbnd;        // [nb][dim*dim + dim] viscous records of the ghost side of each boundary face
    R* gbnd_out;          // written by K1 (vrec)
    const R* qf;          // [ns][n][nv] own-side face states (second order), one record per slot
    R* qf_out;            // written by K1 when not null
    R* qf_mono;           // MONOTONIC1: the face states limited in pairs (k_mono1 from qf_out), which K2
                          // then reads in place of qf; null otherwise
    double* res_out;      // [nv][n] residual-only mode
    R* dinv_out;          // [nv*nv][n] inverse diagonal blocks of the implicit operator (residual mode)
    double* diag_out;     // [nv*nv][n] the diagonal blocks themselves (optional)
    const double* d_dt;   // device time step
    const double* dtloc;  // [n] per-cell time steps (loc_stab_cond) in place of *d_dt, or null
    double* rnorm;        // [n] last stage: update_field's steady residual term of each cell, or null
    double rref[3];       // refsca(rho), refsca(rhoE), refvec: the norm's reference values
    const int* stop;      // steady cycle: nothing to do once *stop is set (null: never)
    // assembled implicit matrix (residual mode with Jacobians; null otherwise): per slot the block
    // coupling this cell's row to the slot's unknown [slot][NV*NV][n], that unknown's index (a cell,
    // or n + ghost: boundary face b -> b, periodic face j -> nb + 2 j (+1 on its far side)), and the
    // ghost rows' right-hand side (their boundary-flux sums) at bdlu_grhs[v * bdlu_ld + ghost]
    double* bdlu_B;
    int* bdlu_tgt;
    double* bdlu_grhs;
    long bdlu_ld;
    int nb_dev, per_base;
    // RK bookkeeping for this stage
    int mode;             // 0: stage update, 1: residual only
    int stage, nstage;
    double a[MAX_STAGES];     // coef(stage, 1..stage) for the next stage state
    double b;                 // coef(nstage, stage)
    int store_slot;           // store R_stage into rstore[store_slot] (-1: no)
    int read_slot[MAX_STAGES];// rstore slot of R_j, j < stage (-1: not stored / zero coefficient)
};

// calc_varprim_ns (SOURCE/EQNS/calc_varprim_ns.f90:47-52) in FP64 on one conservative state
// u = (rho, rhoE, rho u..): the stage-state record (rho, p, u..) in the arithmetic precision R.
// Every producer of stage records (the RK stage update, the state upload, the perturbed states
// of the Krylov products) recovers pressure from total energy here, in FP64, once per cell.
template <typename R>
__device__ __forceinline__ void store_prim(const double* u, int nv, double g1, R* out)
{
    const double rho = u[0];
    const double ir = 1.0 / rho;
    double sq = 0.0;
    for (int d = 0; d < nv - 2; ++d) {
        const double vd = u[2 + d] * ir;
        sq = d == 0 ? vd * vd : sq + vd * vd;
        out[2 + d] = (R)vd;
    }
    out[0] = (R)rho;
    out[1] = (R)(g1 * (u[1] - .5 * rho * sq));
}

struct DtArgs {
    int n, nblk_real;
    const double* u0;
    const double* vol;    // [n]
    const double* sumA;   // [n] sum of the cell's face areas
    const double* cfl;    // device CFL number (fixed, or the steady ramp's)
    double gamma;
    double* part;  // [nblocks]
    int* bad;      // [nblocks]
    double* dtloc; // [n] per-cell time steps are written here when not null (loc_stab_cond)
    const int* stop;  // steady cycle: the kernels do nothing once *stop is set (null: never)
};

// Device state of a steady cycle (integration_cyclezone.f90:80-167 for time_steady): the residual of
// the last iteration, the zone's reference residual, the CFL of the next iteration, counters.
struct SteadyDev {
    double cur_res, residu_ref, cfl;
    int iter, stop;
};
struct SteadyArgs {
    double stabnb, cfl_max, maxres, wlog;
    int maxit, nblk;
    SteadyDev* st;
    const double* part;  // [nblk] block sums of the per-cell residual terms
    double* hist;        // [maxit] cur_res of every iteration
};
// per-block sums of the per-cell steady residual terms, then the cycle bookkeeping (one block)
hipError_t launch_steady_update(const double* rnorm, int n, const SteadyArgs& a, double* part, hipStream_t s);

// launchers (ty_kernels.hip)
template <int DIM, typename R>
hipError_t launch_gradient(const KArgs<R>& a, hipStream_t s);
template <int DIM, typename R>
hipError_t launch_flux(const KArgs<R>& a, hipStream_t s);
template <int DIM>
hipError_t launch_timestep(const DtArgs& a, double dtmax, double* d_dt, int* d_err, hipStream_t s);
// stage-state primitive records [n][nv] from a [nv][n] FP64 conservative state
template <typename R>
hipError_t launch_to_stage(const double* src, R* dst, int nv, int n, double gamma, hipStream_t s);

int grid_blocks(int n);  // XCD-padded block count

// Device inputs of the per-slot geometry build (ty_setup.hip): mesh arrays in FP64 as the
// host renumbered them ([3][n] cell centres, [3][nf] face centres and normals, [nf] areas,
// [3][nb] ghost centres, [3][nper] periodic translations, [n] volumes) and the slot lists.
struct GeoSrc {
    int n, nf, nb, nper, per_base, ns, dim, wlsq;
    double K;
    const double *cc, *fc, *fn, *fs, *gc, *ps, *vol;
    const int *nbr, *fw;
};
// writes ss | sw | sr | sd ([ns][3][n] each) | fn ([3][nf]) | fs ([nf]) | eps2 ([n]) | sh ([ns][n]) in precision R
template <typename R>
hipError_t launch_geo(const GeoSrc& s, R* out, hipStream_t st);
// the FP32 packed slot records of Geo<float> (pa, pb, pw: [ns][n] each in `out`) from its SoA arrays
template <typename R>
hipError_t launch_geo_pack(int n, int ns, const Geo<R>& g, const int* link, typename Pack4<R>::S* out, hipStream_t st);
// stable ascending device sort of n host keys over bits [0, end_bit): idx_out (host) gets the
// input position of each key in sorted order (ty_setup.hip)
hipError_t sort_keys_device(const uint64_t* keys, int n, int end_bit, int* idx_out, hipStream_t st);
// reverse Cuthill-McKee order of an n-node graph given by its edges (ea[e], eb[e]): order[i] = the node
// placed i-th (ty_order.hip); degree_ok false (and no order) if a node has 4096 edges or more
hipError_t rcm_order_device(int n, const std::vector<int>& ea, const std::vector<int>& eb, std::vector<int>& order,
                            bool& degree_ok, hipStream_t st);

}  // namespace ty

// Host side of the C-ABI declared in include/typhon_gpu.h.
//
// Mesh upload turns Typhon's st_ustmesh (ghost cells, periodic boundaries as
// ghost copies, CFDTOOLS/Mesh/USTMESH.f90:31-52) into the device layout:
//   - interior cells renumbered by reverse Cuthill-McKee,
//   - periodic face pairs merged into one two-sided face whose far side is
//     translated by the pair's face-centre difference (init_connect_grid.f90:226-231),
//   - internal faces sorted by their lower cell, boundary faces last,
//   - per-cell face slots in ELL form [slot][cell].
// Everything stays resident; only dt and optional monitors come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/typhon_gpu.h"
#include "ty_internal.h"
#include "ty_krylov.h"
#include "ty_adjoint.h"
#include "ty_probe.h"

using namespace ty;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct ty_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // host mesh (device order)
    int dim = 0, n = 0, ns = 0, nf = 0, nfi = 0, nb = 0;
    int dense = 0;  // every cell has exactly ns face slots
    std::vector<int> perm, iperm;  // perm[new] = caller interior index
    std::vector<int> nbr, fw;      // [ns][n]
    std::vector<int> link;         // [ns][n] link words of the residual kernels
    std::vector<unsigned> bnd_bits;  // bit g of word g / 32: a cell of [32 g, 32 g + 32) has a boundary slot
    std::vector<int> fcell;        // [2][nf] owner and other cell of each device face (-1: boundary)
    std::vector<double> fn, fc, fs;  // [3][nf], [3][nf], [nf]
    std::vector<int> fsrc;           // [nf] caller face (0-based) whose geometry device face k carries
    int nface_caller = 0;            // faces of the caller's mesh
    std::vector<double> cc, vol, sumA;  // [3][n], [n], [n]
    std::vector<double> gc;  // [3][nb]
    std::vector<int> bboco;  // [nb]
    std::vector<double> ps;  // [3][nper] periodic translations, faces per_base..per_base+nper-1
    int per_base = 0, nper = 0;
    std::vector<ty_boco_desc> bocos;  // types and params (pointers not kept)
    bool mesh_ok = false;
    // scheme
    bool scheme_ok = false;
    ty_gas gas{};
    ty_scheme sch{};
    std::vector<double> rk;  // coef(i,j) row-major [ns][ns]
    int nv = 0;
    bool mixed = false;
    bool acc32 = false;  // TY_PREC_MIXED_F32SUM
    Physics ph{};
    // device buffers
    std::vector<void*> owned;
    std::vector<void*> mesh_bufs;  // the mesh upload's buffers among owned (kept by ty_set_scheme)
    int* d_nbr = nullptr;
    int* d_fw = nullptr;
    int* d_link = nullptr;
    unsigned* d_bnd_bits = nullptr;
    int* d_fcell = nullptr;
    void* d_geo = nullptr;  // one allocation for the geometry of the active precision
    Geo<double> geo_d{};
    Geo<float> geo_f{};
    int* d_bboco = nullptr;
    BocoDev* d_boco = nullptr;
    double* d_vol = nullptr;
    // FP64 mesh arrays the device geometry build reads (kept with the mesh)
    double *d_cc = nullptr, *d_fc = nullptr, *d_fn = nullptr, *d_fs = nullptr, *d_gc = nullptr, *d_ps = nullptr;
    std::vector<int> bcell;     // cells with a physical boundary slot, ascending
    int* d_bcell = nullptr;
    std::vector<unsigned char> fslot;  // [2][nf] slot of each device face in its owner's / other cell's list
    unsigned char* d_fslot = nullptr;
    double* d_sumA = nullptr;  // [n] sum of each cell's face areas (mesh)
    double* d_cfl = nullptr;   // the scheme's CFL number
    // steady cycle (ty_steady_cycle): per-cell time steps and residual terms, the cycle's device
    // state; the kernels read them through make_args while a cycle runs
    double *d_dtloc = nullptr, *d_rnorm = nullptr, *d_spart = nullptr, *d_hist = nullptr;
    SteadyDev* d_steady = nullptr;
    int hist_len = 0;
    // volume probes: compiled programs (kept for the context's life) and the device partials +
    // results of ty_probe_eval
    std::vector<ProbeProg> probes;
    double* d_probe = nullptr;
    int probe_cap = 0;
    double* d_scratch = nullptr;  // this device's fixed probe workspace (ty_probe_eval_noalloc)
    bool steady_on = false, steady_local = false;
    double steady_ref[3] = {1.0, 1.0, 1.0};
    double* d_u0 = nullptr;
    void* d_us[2] = {nullptr, nullptr};
    int cur = 0;
    double* d_acc = nullptr;
    double* d_rstore = nullptr;
    int nstore = 0;
    void* d_grad = nullptr;   // work set of the scheme precision (aliases work_grad/qf)
    void* d_qf = nullptr;
    void* work_grad[2] = {nullptr, nullptr};  // per precision: [0] FP64, [1] FP32
    void* work_qf[2] = {nullptr, nullptr};
    void* work_qf2[2] = {nullptr, nullptr};  // MONOTONIC1: the pair-limited face states
    void* work_gbnd[2] = {nullptr, nullptr};  // ghost-side viscous records of the boundary faces
    // matrix-free implicit solve (allocated on first use)
    bool kry_ok = false;
    bool kry_base = false;  // the solver's fixed-size vectors (k_b .. k_h) are allocated
    int kry_m = 0, kry_prec = -1;
    double *k_b = nullptr, *k_x = nullptr, *k_r = nullptr, *k_r0 = nullptr, *k_r1 = nullptr, *k_r2 = nullptr;
    double *k_pd = nullptr, *k_part = nullptr, *k_h = nullptr, *k_dx = nullptr;
    float* k_pf = nullptr;
    void *k_V = nullptr, *k_z = nullptr, *k_w = nullptr, *k_t = nullptr, *k_dinv = nullptr;
    double* d_res = nullptr;
    double* d_part = nullptr;
    double* d_adj_fgeo = nullptr;  // per-face viscous geometry of the adjoint ([4][nf], built on first use)
    int* d_bad = nullptr;
    double* d_dt = nullptr;
    int* d_err = nullptr;
    bool state_ok = false;
    bool ext_bc = false;  // wall or subsonic-inlet boundaries (no tangent/adjoint for them)
    double h_dt = 0.0;
    // per-kernel event timing (bench)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_grad, ev_flux, ev_dt, ev_tan, ev_adj;
};

namespace {

int fail(ty_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

// Setup phase durations on stderr when TY_SETUP_TIMING is set in the environment.
struct PhaseTimer {
    bool on = getenv("TY_SETUP_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void operator()(const char* what)
    {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "# setup %-28s %8.3f s\n", what, std::chrono::duration<double>(now - t).count());
        t = now;
    }
};

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, TY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void free_state(ty_ctx* c)
{
    // cleanup is best effort: a failed free or a sticky error from an earlier launch must not stop it
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->owned) (void)hipFree(p);
    c->owned.clear();
    c->mesh_bufs.clear();
    c->d_nbr = nullptr;
    c->d_fw = nullptr;
    c->d_link = nullptr;
    c->d_bnd_bits = nullptr;
    c->d_fcell = nullptr;
    c->d_geo = nullptr;
    c->d_bboco = nullptr;
    c->d_boco = nullptr;
    c->d_vol = c->d_sumA = c->d_cfl = c->d_u0 = c->d_acc = c->d_rstore = c->d_res = c->d_part = c->d_dt = nullptr;
    c->d_dtloc = c->d_rnorm = c->d_spart = c->d_hist = nullptr;
    c->d_steady = nullptr;
    c->hist_len = 0;
    c->d_probe = nullptr;
    c->probe_cap = 0;
    c->d_cc = c->d_fc = c->d_fn = c->d_fs = c->d_gc = c->d_ps = nullptr;
    c->d_bcell = nullptr;
    c->d_fslot = nullptr;
    c->d_us[0] = c->d_us[1] = nullptr;
    c->d_grad = c->d_qf = nullptr;
    c->work_grad[0] = c->work_grad[1] = c->work_qf[0] = c->work_qf[1] = nullptr;
    c->work_qf2[0] = c->work_qf2[1] = nullptr;
    c->work_gbnd[0] = c->work_gbnd[1] = nullptr;
    c->kry_ok = false;
    c->kry_base = false;
    c->kry_m = 0;
    c->kry_prec = -1;
    c->d_bad = c->d_err = nullptr;
    c->d_adj_fgeo = nullptr;
    for (auto* v : {&c->ev_grad, &c->ev_flux, &c->ev_dt, &c->ev_tan, &c->ev_adj}) {
        for (auto& e : *v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        v->clear();
    }
}

template <typename T>
int dalloc(ty_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(c, TY_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->owned.push_back(q);
    *p = static_cast<T*>(q);
    return TY_OK;
}

template <typename T>
int upload(ty_ctx* c, T** p, const std::vector<T>& h)
{
    int rc = dalloc(c, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIPCHK(c, hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return TY_OK;
}

// fn(i) for i in [0, n) over up to 16 host threads (contiguous chunks); serial for small n
template <typename Fn>
void parallel_for(int n, Fn fn)
{
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < (1 << 16) || nt == 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&fn, lo, hi] {
            for (int i = lo; i < hi; ++i) fn(i);
        });
    }
    for (auto& x : th) x.join();
}

// Hilbert-curve order of the cell centres (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707 (2004) 381: axes -> transposed index, then bit interleave)
int hilbert_order(ty_ctx* c, int n, const double* centre, int dim, std::vector<int>& order)
{
    const int bits = dim == 3 ? 21 : 31;
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) {
            lo[d] = std::min(lo[d], centre[(size_t)3 * i + d]);
            hi[d] = std::max(hi[d], centre[(size_t)3 * i + d]);
        }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], dim == 3 ? hi[2] - lo[2] : 0.0, 1e-300});
    const double scale = (double)((1u << (bits - 1)) - 1) * 2.0 / span;
    std::vector<uint64_t> key(n);
    parallel_for(n, [&](int i) {
        uint32_t X[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) X[d] = (uint32_t)((centre[(size_t)3 * i + d] - lo[d]) * scale);
        const uint32_t M = 1u << (bits - 1);
        for (uint32_t Q = M; Q > 1; Q >>= 1) {
            const uint32_t P = Q - 1;
            for (int d = 0; d < dim; ++d) {
                if (X[d] & Q) X[0] ^= P;
                else { uint32_t t = (X[0] ^ X[d]) & P; X[0] ^= t; X[d] ^= t; }
            }
        }
        for (int d = 1; d < dim; ++d) X[d] ^= X[d - 1];
        uint32_t t = 0;
        for (uint32_t Q = M; Q > 1; Q >>= 1)
            if (X[dim - 1] & Q) t ^= Q - 1;
        for (int d = 0; d < dim; ++d) X[d] ^= t;
        uint64_t h = 0;
        for (int b = bits - 1; b >= 0; --b)
            for (int d = 0; d < dim; ++d) h = (h << 1) | ((X[d] >> b) & 1u);
        key[i] = h;
    });
    // stable in the cell index: the order std::stable_sort gives on (key, cell)
    order.resize(n);
    HIPCHK(c, sort_keys_device(key.data(), n, 64, order.data(), c->stream));
    return TY_OK;
}

}  // namespace

// The C-ABI functions below take C linkage from their declarations in typhon_gpu.h.

int ty_create(ty_ctx** ctx, int device)
{
    if (!ctx) return TY_ERR_ARG;
    *ctx = nullptr;
    ty_ctx* c = new ty_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = probe_scratch_setup(device, &c->d_scratch);
    if (e != hipSuccess) {
        delete c;
        return TY_ERR_HIP;
    }
    *ctx = c;
    return TY_OK;
}

int ty_destroy(ty_ctx* c)
{
    if (!c) return TY_OK;
    (void)hipSetDevice(c->device);  // best effort: destruction frees what it can
    free_state(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TY_OK;
}

const char* ty_last_error(const ty_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ty_mesh_upload(ty_ctx* c, const ty_mesh_desc* m)
{
    if (!c || !m) return TY_ERR_ARG;
    if (m->dim != 2 && m->dim != 3) return fail(c, TY_ERR_ARG, "dim must be 2 or 3");
    const int ni = m->ncell_int, nc = m->ncell, nf = m->nface, nfi = m->nface_int;
    if (ni <= 0 || nc < ni || nf <= 0 || nfi < 0 || nfi > nf) return fail(c, TY_ERR_ARG, "bad mesh sizes");
    HIPCHK(c, hipSetDevice(c->device));
    PhaseTimer phase;
    free_state(c);
    c->mesh_ok = c->scheme_ok = c->state_ok = false;
    c->dim = m->dim;
    c->n = ni;
    std::vector<int> ic1(nf), ic2(nf);
    for (int f = 0; f < nf; ++f) {
        ic1[f] = m->facecell[f] - 1;
        ic2[f] = m->facecell[(size_t)nf + f] - 1;
        if (ic1[f] < 0 || ic1[f] >= ni) return fail(c, TY_ERR_ARG, "face " + std::to_string(f + 1) + ": bad ic1");
        if (f < nfi && (ic2[f] < 0 || ic2[f] >= ni))
            return fail(c, TY_ERR_ARG, "internal face " + std::to_string(f + 1) + ": bad ic2");
        if (f >= nfi && (ic2[f] < ni || ic2[f] >= nc))
            return fail(c, TY_ERR_ARG, "boundary face " + std::to_string(f + 1) + ": ic2 is not a ghost cell");
    }
    auto F = [&](int f, int k) { return m->face[(size_t)7 * f + k]; };
    // boundary faces -> (boco, partner)
    std::vector<int> fboco(nf, -1), fpart(nf, -1);
    c->bocos.assign(m->boco, m->boco + m->nboco);
    if (m->nboco > MAX_BOCO) return fail(c, TY_ERR_UNSUPPORTED, "too many boundary conditions");
    for (int b = 0; b < m->nboco; ++b) {
        const ty_boco_desc& d = m->boco[b];
        c->bocos[b].iface = nullptr;
        c->bocos[b].ipart = nullptr;
        for (int k = 0; k < d.nface; ++k) {
            int f = d.iface[k] - 1;
            if (f < nfi || f >= nf) return fail(c, TY_ERR_ARG, "boco face outside the boundary face range");
            if (fboco[f] >= 0) return fail(c, TY_ERR_ARG, "face in two boundary conditions");
            fboco[f] = b;
            if (d.type == 11) {
                if (!d.ipart) return fail(c, TY_ERR_ARG, "periodic boco without partner faces");
                int p = d.ipart[k] - 1;
                if (p < nfi || p >= nf) return fail(c, TY_ERR_ARG, "bad periodic partner face");
                fpart[f] = p;
            } else if (d.type == 23 || d.type == 24) {
                // convection walls: the reference refuses them for NS at set-up (bctype_of_nsboco,
                // MENU_NS.f90:95-113, has no case for them) and setboco_ns_hconv.f90:78 stops anyway
                return fail(c, TY_ERR_UNSUPPORTED, "boundary condition type " + std::to_string(d.type) +
                                                       " (convection wall): the reference stops in "
                                                       "setboco_ns_hconv (not implemented there)");
            } else if (d.type != 10 && d.type != 20 && d.type != 21 && d.type != 22 && d.type != 30 &&
                       d.type != 31 && d.type != 35 && d.type != 36) {
                return fail(c, TY_ERR_UNSUPPORTED, "boundary condition type " + std::to_string(d.type) +
                                                       " is not provided by this library");
            }
        }
    }
    for (int f = nfi; f < nf; ++f)
        if (fboco[f] < 0) return fail(c, TY_ERR_ARG, "boundary face " + std::to_string(f + 1) + " has no boco");
    for (int f = nfi; f < nf; ++f)
        if (fpart[f] >= 0 && fpart[fpart[f]] != f) return fail(c, TY_ERR_ARG, "periodic partners are not mutual");
    // device faces before renumbering: owner, other (or -1 boundary), source face, periodic
    // partner face (or -1), boco index
    struct DF { int a, b, src, part, bidx; };
    std::vector<DF> faces, pfaces, bfaces;
    faces.reserve(nfi);
    for (int f = 0; f < nfi; ++f) faces.push_back({ic1[f], ic2[f], f, -1, -1});
    for (int f = nfi; f < nf; ++f) {
        if (fpart[f] >= 0) {
            int p = fpart[f];
            if (f < p) pfaces.push_back({ic1[f], ic1[p], f, p, -1});
        } else {
            bfaces.push_back({ic1[f], -1, f, -1, fboco[f]});
        }
    }
    phase("mesh: checks, face lists");
    // cell order: RCM over the face graph, a Hilbert curve through the cell centres, or the caller's
    std::vector<int> order;
    if (m->cell_order == 1) {
        int rc0 = hilbert_order(c, ni, m->centre, m->dim, order);
        if (rc0) return rc0;
    } else if (m->cell_order == 2) {
        order.resize(ni);
        for (int i = 0; i < ni; ++i) order[i] = i;
    } else {
        // reverse Cuthill-McKee over the graph of interior faces and merged periodic pairs, on the
        // device; adjacency lists in this edge order
        std::vector<int> ea, eb;
        ea.reserve(faces.size() + pfaces.size());
        eb.reserve(faces.size() + pfaces.size());
        for (auto* L : {&faces, &pfaces})
            for (auto& d : *L) {
                ea.push_back(d.a);
                eb.push_back(d.b);
            }
        bool degree_ok = true;
        HIPCHK(c, rcm_order_device(ni, ea, eb, order, degree_ok, c->stream));
        if (!degree_ok) return fail(c, TY_ERR_UNSUPPORTED, "reverse Cuthill-McKee order: a cell with 4096 neighbours or more");
    }
    phase("mesh: cell order");
    c->perm = order;
    c->iperm.assign(ni, 0);
    for (int i = 0; i < ni; ++i) c->iperm[c->perm[i]] = i;
    for (auto* L : {&faces, &pfaces})
        parallel_for((int)L->size(), [&](int k) { DF& d = (*L)[k]; d.a = c->iperm[d.a]; d.b = c->iperm[d.b]; });
    parallel_for((int)bfaces.size(), [&](int k) { bfaces[k].a = c->iperm[bfaces[k].a]; });
    // faces ordered by (lower cell, higher cell), boundary faces by cell, ties in list order:
    // a stable device radix sort of the packed keys
    auto sort_faces = [&](std::vector<DF>& L, bool two) -> int {
        const int mf = (int)L.size();
        std::vector<uint64_t> key(mf);
        parallel_for(mf, [&](int k) {
            const DF& d = L[k];
            key[k] = two ? ((uint64_t)(uint32_t)std::min(d.a, d.b) << 32) | (uint32_t)std::max(d.a, d.b)
                         : (uint64_t)(uint32_t)d.a;
        });
        std::vector<int> idx(mf);
        HIPCHK(c, sort_keys_device(key.data(), mf, two ? 64 : 32, idx.data(), c->stream));
        std::vector<DF> out(mf);
        parallel_for(mf, [&](int k) { out[k] = L[idx[k]]; });
        L.swap(out);
        return TY_OK;
    };
    {
        int rc0;
        if ((rc0 = sort_faces(faces, true)) || (rc0 = sort_faces(pfaces, true)) || (rc0 = sort_faces(bfaces, false)))
            return rc0;
    }
    phase("mesh: face renumbering");
    // device face ranges: internal | periodic pairs | physical boundary
    c->per_base = (int)faces.size();
    c->nper = (int)pfaces.size();
    c->nfi = c->per_base + c->nper;
    c->nb = (int)bfaces.size();
    c->nf = c->nfi + c->nb;
    const int n = ni, ndf = c->nf;
    if ((unsigned)ndf > FACE_MASK) return fail(c, TY_ERR_UNSUPPORTED, "too many faces for the slot word");
    c->fn.assign((size_t)3 * ndf, 0.0);
    c->fc.assign((size_t)3 * ndf, 0.0);
    c->fs.assign(ndf, 0.0);
    c->fsrc.assign(ndf, -1);
    c->nface_caller = nf;
    c->gc.assign((size_t)3 * c->nb, 0.0);
    c->bboco.assign(c->nb, 0);
    c->ps.assign((size_t)3 * std::max(1, c->nper), 0.0);
    auto put_face = [&](int df, int src) {
        for (int d = 0; d < 3; ++d) {
            c->fn[(size_t)d * ndf + df] = F(src, d);
            c->fc[(size_t)d * ndf + df] = F(src, 3 + d);
        }
        c->fs[df] = F(src, 6);
        c->fsrc[df] = src;
    };
    c->fcell.assign((size_t)2 * ndf, -1);
    auto dface = [&](int k) -> const DF& { return k >= c->per_base ? pfaces[k - c->per_base] : faces[k]; };
    parallel_for(c->nfi, [&](int k) {
        const DF& d = dface(k);
        put_face(k, d.src);
        c->fcell[k] = d.a;
        c->fcell[(size_t)ndf + k] = d.b;
        if (k >= c->per_base) {
            // far side seen through this face: x_far + (c_face - c_partner) (init_connect_grid.f90:226-231)
            const int j = k - c->per_base;
            for (int e = 0; e < 3; ++e) c->ps[(size_t)e * c->nper + j] = F(d.src, 3 + e) - F(d.part, 3 + e);
        }
    });
    parallel_for(c->nb, [&](int k) {
        const DF& d = bfaces[k];
        const int df = c->nfi + k;
        put_face(df, d.src);
        c->fcell[df] = d.a;
        const int g = ic2[d.src];
        for (int e = 0; e < 3; ++e) c->gc[(size_t)e * c->nb + k] = m->centre[(size_t)3 * g + e];
        c->bboco[k] = d.bidx;
    });
    phase("mesh: faces, slot lists");
    // slot lists (CSR): each cell's faces in ascending device face order, as (neighbour or
    // -(boundary face + 1), slot word = face | periodic bit | other-side bit)
    std::vector<int> sbeg(n + 1, 0);
    for (int k = 0; k < c->nfi; ++k) { const DF& d = dface(k); ++sbeg[d.a + 1]; ++sbeg[d.b + 1]; }
    for (int k = 0; k < c->nb; ++k) ++sbeg[bfaces[k].a + 1];
    int ns = 0;
    c->dense = 1;
    for (int i = 0; i < n; ++i) ns = std::max(ns, sbeg[i + 1]);
    for (int i = 0; i < n; ++i)
        if (sbeg[i + 1] != ns) c->dense = 0;
    for (int i = 0; i < n; ++i) sbeg[i + 1] += sbeg[i];
    if (ns > MAX_SLOTS) return fail(c, TY_ERR_UNSUPPORTED, "more than 8 faces per cell");
    c->ns = ns;
    std::vector<int> snb(sbeg[n]), sword(sbeg[n]), spos(sbeg.begin(), sbeg.end() - 1);
    for (int k = 0; k < c->nfi; ++k) {
        const DF& d = dface(k);
        const unsigned w = (unsigned)k | (k >= c->per_base ? PERIODIC_BIT : 0u);
        snb[spos[d.a]] = d.b;
        sword[spos[d.a]++] = (int)w;
        snb[spos[d.b]] = d.a;
        sword[spos[d.b]++] = (int)(w | 0x80000000u);
    }
    for (int k = 0; k < c->nb; ++k) {
        const int a = bfaces[k].a;
        snb[spos[a]] = -(k + 1);
        sword[spos[a]++] = c->nfi + k;
    }
    c->nbr.assign((size_t)ns * n, 0);
    c->fw.assign((size_t)ns * n, SLOT_UNUSED);
    c->link.assign((size_t)ns * n, SLOT_UNUSED);
    if (n >= (1 << 27) || c->nb >= (1 << 27)) return fail(c, TY_ERR_UNSUPPORTED, "too many cells for the slot link word");
    c->sumA.assign(n, 0.0);
    // the slot where each side keeps each device face (each face side is written once): the
    // mirror slots of the slot words and the adjoint's per-side record positions
    c->fslot.assign((size_t)2 * ndf, 0);
    unsigned char* slot_own = c->fslot.data();
    unsigned char* slot_oth = c->fslot.data() + ndf;
    parallel_for(n, [&](int i) {
        for (int k = 0; k < sbeg[i + 1] - sbeg[i]; ++k) {
            const unsigned w = (unsigned)sword[sbeg[i] + k];
            (w & 0x80000000u ? slot_oth : slot_own)[w & FACE_MASK] = (unsigned char)k;
        }
    });
    parallel_for(n, [&](int i) {
        for (int k = 0; k < sbeg[i + 1] - sbeg[i]; ++k) {
            const int nbv = snb[sbeg[i] + k];
            unsigned w = (unsigned)sword[sbeg[i] + k];
            if (nbv >= 0) {
                const unsigned mk = (w & 0x80000000u) ? slot_own[w & FACE_MASK] : slot_oth[w & FACE_MASK];
                w |= mk << MIRROR_SHIFT;
            }
            c->nbr[(size_t)k * n + i] = nbv;
            c->fw[(size_t)k * n + i] = (int)w;
            if (nbv >= 0) {
                const unsigned mk = (w >> MIRROR_SHIFT) & (MAX_SLOTS - 1);
                c->link[(size_t)k * n + i] = (int)(((w & 0x80000000u) ? (1u << 30) : 0u) | (mk << 27) | (unsigned)nbv);
            } else {
                c->link[(size_t)k * n + i] = (int)(0x80000000u | (unsigned)(-nbv - 1));
            }
            c->sumA[i] += c->fs[w & FACE_MASK];
        }
    });
    c->cc.assign((size_t)3 * n, 0.0);
    c->vol.assign(n, 0.0);
    parallel_for(n, [&](int i) {
        const int o = c->perm[i];
        for (int d = 0; d < 3; ++d) c->cc[(size_t)d * n + i] = m->centre[(size_t)3 * o + d];
        c->vol[i] = m->volume[o];
    });
    // groups of 32 cells with a boundary slot: the residual kernels take a path without the boundary
    // code for the wavefronts whose cells have none
    c->bnd_bits.assign(((size_t)n + 1023) / 1024, 0u);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < c->ns; ++k) {
            const int L = c->link[(size_t)k * n + i];
            if (L != SLOT_UNUSED && L < 0) {
                c->bnd_bits[(size_t)i >> 10] |= 1u << ((i >> 5) & 31);
                break;
            }
        }
    }
    c->bcell.clear();
    for (int k = 0; k < c->nb; ++k) c->bcell.push_back(bfaces[k].a);  // bfaces are sorted by cell
    c->bcell.erase(std::unique(c->bcell.begin(), c->bcell.end()), c->bcell.end());
    phase("mesh: slot words, cell data");
    int rc;
    if ((rc = upload(c, &c->d_nbr, c->nbr)) || (rc = upload(c, &c->d_fw, c->fw)) ||
        (rc = upload(c, &c->d_bboco, c->bboco)) || (rc = upload(c, &c->d_vol, c->vol)) ||
        (rc = upload(c, &c->d_link, c->link)) || (rc = upload(c, &c->d_bnd_bits, c->bnd_bits)) ||
        (rc = upload(c, &c->d_fcell, c->fcell)) || (rc = upload(c, &c->d_cc, c->cc)) ||
        (rc = upload(c, &c->d_fc, c->fc)) || (rc = upload(c, &c->d_fn, c->fn)) || (rc = upload(c, &c->d_fs, c->fs)) ||
        (rc = upload(c, &c->d_gc, c->gc)) || (rc = upload(c, &c->d_ps, c->ps)) ||
        (rc = upload(c, &c->d_bcell, c->bcell)) || (rc = upload(c, &c->d_fslot, c->fslot)) ||
        (rc = upload(c, &c->d_sumA, c->sumA)))
        return rc;
    phase("mesh: uploads");
    c->mesh_bufs = c->owned;  // everything allocated so far belongs to the mesh
    c->mesh_ok = true;
    return TY_OK;
}

int ty_mesh_info(const ty_ctx* c, int32_t* ncell, int32_t* nface_dev, int32_t* nslot, int32_t* nbface_dev)
{
    if (!c || !c->mesh_ok) return TY_ERR_ARG;
    if (ncell) *ncell = c->n;
    if (nface_dev) *nface_dev = c->nf;
    if (nslot) *nslot = c->ns;
    if (nbface_dev) *nbface_dev = c->nb;
    return TY_OK;
}

int ty_cell_order(const ty_ctx* c, int32_t* cell)
{
    if (!c || !c->mesh_ok || !cell) return TY_ERR_ARG;
    for (int i = 0; i < c->n; ++i) cell[i] = c->perm[i] + 1;
    return TY_OK;
}

namespace {

template <typename R>
int build_geo(ty_ctx* c, Geo<R>& g)
{
    // per-slot geometry, built on the device from the mesh arrays (ty_setup.hip):
    // ss | sw | sr | sd ([ns][3][n] each) | fn ([3][nf]) | fs ([nf]) | eps2 ([n]) | sh ([ns][n])
    const int n = c->n, nf = c->nf, ns = c->ns;
    const size_t S3 = (size_t)3 * ns * n;
    const size_t o_ss = 0, o_sw = o_ss + S3, o_sr = o_sw + S3, o_sd = o_sr + S3;
    const size_t o_fn = o_sd + S3, o_fs = o_fn + 3 * (size_t)nf, o_eps = o_fs + nf, o_sh = o_eps + n;
    const size_t tot = o_sh + (size_t)ns * n;
    R* dp = nullptr;
    int rc = dalloc(c, &dp, tot);
    if (rc) return rc;
    GeoSrc src;
    src.n = n;
    src.nf = nf;
    src.nb = c->nb;
    src.nper = std::max(1, c->nper);
    src.per_base = c->per_base;
    src.ns = ns;
    src.dim = c->dim;
    src.wlsq = c->sch.gradmeth == 12;
    src.K = c->sch.venkat_k;
    src.cc = c->d_cc;
    src.fc = c->d_fc;
    src.fn = c->d_fn;
    src.fs = c->d_fs;
    src.gc = c->d_gc;
    src.ps = c->d_ps;
    src.vol = c->d_vol;
    src.nbr = c->d_nbr;
    src.fw = c->d_fw;
    HIPCHK(c, launch_geo<R>(src, dp, c->stream));
    g.ss = dp + o_ss;
    g.sw = dp + o_sw;
    g.sr = dp + o_sr;
    g.sd = dp + o_sd;
    g.fnx = dp + o_fn;
    g.fny = dp + o_fn + nf;
    g.fnz = dp + o_fn + 2 * (size_t)nf;
    g.fs = dp + o_fs;
    g.eps2 = dp + o_eps;
    g.sh = dp + o_sh;
    g.bboco = c->d_bboco;
    // the packed slot records of the interior paths (viscous and plain-gradient kernels)
    // (FP32: pw, pa, pb; FP64: pw, the FP64 flux kernel keeps its scalar loads)
    typename Pack4<R>::S* pk = nullptr;
    const size_t arr = (size_t)Pack4<R>::planes * ns * n;
    const bool fp32 = std::is_same<R, float>::value;
    if ((rc = dalloc(c, &pk, (fp32 ? 3 : 1) * arr))) return rc;
    HIPCHK(c, launch_geo_pack<R>(n, ns, g, c->d_link, pk, c->stream));
    g.pw = pk;
    g.pa = fp32 ? pk + arr : nullptr;
    g.pb = fp32 ? pk + 2 * arr : nullptr;
    return TY_OK;
}

}  // namespace

namespace {

// gradient / face-state work arrays of one arithmetic precision (0: FP64, 1: FP32)
int ensure_work(ty_ctx* c, int prec)
{
    const size_t rb = prec ? sizeof(float) : sizeof(double);
    void* p = nullptr;
    if (c->ph.viscous && !c->work_grad[prec]) {  // cell gradients for the viscous flux
        HIPCHK(c, hipMalloc(&p, (size_t)c->nv * c->dim * c->n * rb));
        c->owned.push_back(p);
        c->work_grad[prec] = p;
    }
    if (c->ph.viscous && !c->work_gbnd[prec]) {  // their ghost sides on boundary faces
        HIPCHK(c, hipMalloc(&p, (size_t)(c->dim * c->dim + c->dim) * std::max(1, c->nb) * rb));
        c->owned.push_back(p);
        c->work_gbnd[prec] = p;
    }
    if (c->ph.method != 'N' && !c->work_qf[prec]) {  // own-side face states of every slot
        HIPCHK(c, hipMalloc(&p, (size_t)c->ns * c->nv * c->n * rb));
        c->owned.push_back(p);
        c->work_qf[prec] = p;
    }
    if (c->ph.mono == 11 && !c->work_qf2[prec]) {
        HIPCHK(c, hipMalloc(&p, (size_t)c->ns * c->nv * c->n * rb));
        c->owned.push_back(p);
        c->work_qf2[prec] = p;
    }
    return TY_OK;
}

}  // namespace

int ty_set_scheme(ty_ctx* c, const ty_gas* gas, const ty_scheme* s)
{
    if (!c || !gas || !s) return TY_ERR_ARG;
    if (!c->mesh_ok) return fail(c, TY_ERR_ARG, "ty_set_scheme before ty_mesh_upload");
    if (s->nstage < 1 || s->nstage > MAX_STAGES || !s->rk_coef) return fail(c, TY_ERR_ARG, "bad RK tableau");
    const int sch = s->sch_hyp;
    if (sch != 5 && sch != 10 && (sch < 21 || sch > 28))
        return fail(c, TY_ERR_UNSUPPORTED, "numerical flux " + std::to_string(sch) + " not provided");
    if (!strchr("NMFU", s->method) || s->method == 0) return fail(c, TY_ERR_UNSUPPORTED, "high-resolution method");
    if (!strchr("XMAVSKLW", s->limiter) || s->limiter == 0) return fail(c, TY_ERR_UNSUPPORTED, "limiter");
    if (s->postlimiter != 0 && s->postlimiter != 20 && s->postlimiter != 21 && (s->postlimiter < 10 || s->postlimiter > 12))
        return fail(c, TY_ERR_UNSUPPORTED, "post-limiter");
    if (s->limiter == 'W' && s->postlimiter != 0)
        return fail(c, TY_ERR_UNSUPPORTED, "Venkatakrishnan limiter with a post-limiter");
    const bool viscous = gas->typ_fluid == 15;
    if (gas->typ_fluid != 10 && !viscous) return fail(c, TY_ERR_UNSUPPORTED, "fluid model");
    // the reference's NS residual runs CELL-AVERAGE (5) and FULL (10) only: COMPACT (2) stops in
    // interp_facegradient_vect (interpface_gradient_vect.f90:61-63, called by
    // calc_flux_viscous_cell.f90:111); FACE-AVERAGE (20) and PENALTY (22) need face gradients, which
    // get_gradientmethod refuses for first-order and MUSCL methods (MENU_NUM.f90:307-309, 318-319)
    if (viscous && s->sch_dis == 2)
        return fail(c, TY_ERR_UNSUPPORTED, "dissipative flux COMPACT: the NS velocity-gradient interpolation "
                                           "does not authorise it (use CELL-AVERAGE or FULL)");
    if (viscous && (s->sch_dis == 20 || s->sch_dis == 22))
        return fail(c, TY_ERR_UNSUPPORTED, "dissipative flux FACE-AVERAGE/PENALTY: face based gradients are "
                                           "unavailable for first-order and MUSCL methods");
    if (viscous && s->sch_dis != 5 && s->sch_dis != 10)
        return fail(c, TY_ERR_UNSUPPORTED, "dissipative flux scheme");
    if ((s->method != 'N' || viscous) && s->gradmeth != 10 && s->gradmeth != 12)
        return fail(c, TY_ERR_UNSUPPORTED, "gradient method (LSQ or W-LSQ required)");
    if (s->precision != TY_PREC_FP64 && s->precision != TY_PREC_MIXED && s->precision != TY_PREC_MIXED_F32SUM)
        return fail(c, TY_ERR_ARG, "precision");
    for (const ty_boco_desc& d : c->bocos)
        if ((d.type == 20 || d.type == 22) && !viscous)
            return fail(c, TY_ERR_UNSUPPORTED, "flux and adiabatic walls need the laminar NS model "
                                               "(their ghost state and ns_bocoflux use the conductivity)");
    HIPCHK(c, hipSetDevice(c->device));
    // From here on the context holds no scheme and no state until the rebuild below completes:
    // a failure part-way leaves scheme_ok / state_ok false and no pointer to a freed buffer.
    c->scheme_ok = c->state_ok = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->d_boco = nullptr;
    c->d_geo = nullptr;
    c->geo_d = Geo<double>{};
    c->geo_f = Geo<float>{};
    c->d_cfl = c->d_u0 = c->d_acc = c->d_rstore = c->d_res = c->d_part = c->d_dt = nullptr;
    c->d_dtloc = c->d_rnorm = c->d_spart = c->d_hist = nullptr;
    c->d_steady = nullptr;
    c->hist_len = 0;
    c->d_probe = nullptr;
    c->probe_cap = 0;
    c->d_us[0] = c->d_us[1] = nullptr;
    c->d_grad = c->d_qf = nullptr;
    c->work_grad[0] = c->work_grad[1] = c->work_qf[0] = c->work_qf[1] = nullptr;
    c->work_qf2[0] = c->work_qf2[1] = nullptr;
    c->work_gbnd[0] = c->work_gbnd[1] = nullptr;
    c->kry_ok = c->kry_base = false;
    c->kry_m = 0;
    c->kry_prec = -1;
    c->d_bad = c->d_err = nullptr;
    c->d_adj_fgeo = nullptr;
    // keep the mesh buffers, drop the rest
    {
        std::vector<void*> keep;
        for (void* p : c->owned) {
            if (std::find(c->mesh_bufs.begin(), c->mesh_bufs.end(), p) != c->mesh_bufs.end()) keep.push_back(p);
            else (void)hipFree(p);  // an earlier scheme's buffers: best effort
        }
        c->owned = keep;
    }
    c->gas = *gas;
    c->sch = *s;
    const int ns = s->nstage;
    c->rk.assign((size_t)ns * ns, 0.0);
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) c->rk[(size_t)i * ns + j] = s->rk_coef[(size_t)j * ns + i];
    c->sch.rk_coef = nullptr;
    c->nv = c->dim + 2;
    c->mixed = s->precision != TY_PREC_FP64;
    c->acc32 = s->precision == TY_PREC_MIXED_F32SUM;
    Physics& ph = c->ph;
    ph.gamma = gas->gamma;
    ph.rgas = gas->r_const;
    ph.prandtl = gas->prandtl;
    ph.visc_law = gas->typ_visc;
    ph.visc_dyn = gas->visc_dyn;
    ph.visc_kin = gas->visc_kin;
    ph.tref = gas->tref;
    ph.tsuth = gas->tsuth;
    ph.viscous = viscous;
    ph.sch_dis = s->sch_dis;
    ph.sch_hyp = sch;
    ph.method = s->method;
    ph.limiter = s->method == 'N' ? 'X' : s->limiter;
    ph.lsqw = s->gradmeth == 12;
    // post-limiters act on second-order face states only (on cell values they change nothing)
    ph.barth = s->method != 'N' && s->limiter != 'W' && (s->postlimiter == 20 || s->postlimiter == 21);
    ph.klim = s->postlimiter == 21 ? 1.0 : 0.5;
    ph.mono = (s->method != 'N' && s->limiter != 'W' && s->postlimiter >= 10 && s->postlimiter <= 12) ? s->postlimiter : 0;
    ph.ig1 = 1.0 / (gas->gamma - 1.0);
    ph.gig1 = gas->gamma / (gas->gamma - 1.0);
    ph.irgas = 1.0 / gas->r_const;
    ph.cp_pr = gas->gamma * gas->r_const / (gas->gamma - 1.0) / gas->prandtl;
    ph.itref = gas->tref != 0.0 ? 1.0 / gas->tref : 0.0;
    ph.tsum = gas->tref + gas->tsuth;
    // boundary table: supersonic inlet state from pi_ti_mach_dir2nspri (EQNS.f90:139-160)
    std::vector<BocoDev> tab(std::max<size_t>(1, c->bocos.size()));
    for (size_t b = 0; b < c->bocos.size(); ++b) {
        const ty_boco_desc& d = c->bocos[b];
        BocoDev& t = tab[b];
        memset(&t, 0, sizeof(t));
        switch (d.type) {
        case 10: t.kind = BK_SYM; break;
        case 35: t.kind = BK_OUTLET_SUB; t.param[0] = d.param[0]; break;
        case 36: t.kind = BK_OUTLET_SUP; break;
        case 31: {
            t.kind = BK_INLET_SUP;
            const double g = gas->gamma, g1 = g - 1.0, r = gas->r_const;
            double dx = d.param[3], dy = d.param[4], dz = d.param[5];
            const double dn = std::sqrt(dx * dx + dy * dy + dz * dz);
            dx /= dn; dy /= dn; dz /= dn;
            const double mach = d.param[2];
            const double fm = 1.0 / (1.0 + .5 * g1 * mach * mach);
            const double ts = d.param[1] * fm;
            const double a = std::sqrt(g * r * ts);
            const double p = d.param[0] * std::pow(fm, g / g1);
            const double rho = p / (r * ts);
            const double ma = mach * a;
            t.param[0] = rho;
            t.param[1] = p;
            t.param[2] = ma * dx;
            t.param[3] = ma * dy;
            if (c->dim == 3) t.param[4] = ma * dz;
            break;
        }
        case 20:  // adiabatic wall: a flux wall with zero flux (def_boco_ns.f90:44-49)
        case 22:
        case 21:
            t.kind = d.type == 21 ? BK_WALL_ISOTH : BK_WALL_FLUX;
            t.param[0] = d.type == 20 ? 0.0 : d.param[0];
            for (int k = 4; k < 7; ++k) t.param[k] = d.param[k];
            break;
        case 30: {
            // uniform PI, TI, DIRECTION; dir/|dir| as setboco_ns_inlet_sub.f90:107
            t.kind = BK_INLET_SUB;
            t.param[0] = d.param[0];
            t.param[1] = d.param[1];
            const double dn = std::sqrt(d.param[3] * d.param[3] + d.param[4] * d.param[4] + d.param[5] * d.param[5]);
            for (int k = 3; k < 6; ++k) t.param[k] = d.param[k] / dn;
            break;
        }
        default: t.kind = BK_OUTLET_SUP;
        }
    }
    int rc;
    if ((rc = upload(c, &c->d_boco, tab))) return rc;
    c->ext_bc = false;
    for (const BocoDev& t : tab)
        if (t.kind == BK_WALL_ISOTH || t.kind == BK_WALL_FLUX || t.kind == BK_INLET_SUB) c->ext_bc = true;
    // both precisions: the FP32 geometry also serves the FP32 Krylov products
    PhaseTimer phase;
    if ((rc = build_geo<float>(c, c->geo_f)) || (rc = build_geo<double>(c, c->geo_d))) return rc;
    if (phase.on) HIPCHK(c, hipStreamSynchronize(c->stream));
    phase("scheme: slot geometry (device)");
    if ((rc = upload(c, &c->d_cfl, std::vector<double>{s->cfl}))) return rc;
    // RK storage plan: R_s is kept if a later non-final stage state uses it off the diagonal
    c->nstore = 0;
    const size_t cells = (size_t)c->nv * c->n;
    int nstore = 0;
    for (int sidx = 1; sidx <= ns; ++sidx)
        for (int t = sidx + 1; t < ns; ++t)
            if (c->rk[(size_t)(t - 1) * ns + (sidx - 1)] != 0.0) { ++nstore; break; }
    c->nstore = nstore;
    const size_t rb = c->mixed ? sizeof(float) : sizeof(double);
    void* p = nullptr;
    if ((rc = dalloc(c, &c->d_u0, cells)) || (rc = dalloc(c, &c->d_acc, cells)) ||
        (rc = dalloc(c, &c->d_rstore, cells * std::max(1, nstore))) || (rc = dalloc(c, &c->d_res, cells)))
        return rc;
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c, hipMalloc(&p, cells * rb));
        c->owned.push_back(p);
        c->d_us[k] = p;
    }
    if ((rc = ensure_work(c, c->mixed ? 1 : 0))) return rc;
    c->d_grad = c->work_grad[c->mixed ? 1 : 0];
    c->d_qf = c->work_qf[c->mixed ? 1 : 0];
    const int nblk = grid_blocks(c->n);
    if ((rc = dalloc(c, &c->d_part, nblk)) || (rc = dalloc(c, &c->d_bad, nblk)) || (rc = dalloc(c, &c->d_dt, 1)) ||
        (rc = dalloc(c, &c->d_err, 1)))
        return rc;
    HIPCHK(c, hipMemset(c->d_err, 0, sizeof(int)));
    c->cur = 0;
    c->scheme_ok = true;
    c->state_ok = false;
    return TY_OK;
}

int ty_state_set(ty_ctx* c, const double* rho, const double* rhoE, const double* mom)
{
    if (!c || !rho || !rhoE || !mom) return TY_ERR_ARG;
    if (!c->scheme_ok) return fail(c, TY_ERR_ARG, "ty_state_set before ty_set_scheme");
    HIPCHK(c, hipSetDevice(c->device));
    const int n = c->n, nv = c->nv;
    std::vector<double> h((size_t)nv * n);
    for (int i = 0; i < n; ++i) {
        int o = c->perm[i];
        h[i] = rho[o];
        h[(size_t)n + i] = rhoE[o];
        for (int d = 0; d < c->dim; ++d) h[(size_t)(2 + d) * n + i] = mom[(size_t)3 * o + d];
    }
    HIPCHK(c, hipMemcpyAsync(c->d_u0, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    hipError_t e = c->mixed ? launch_to_stage<float>(c->d_u0, (float*)c->d_us[0], nv, n, c->gas.gamma, c->stream)
                            : launch_to_stage<double>(c->d_u0, (double*)c->d_us[0], nv, n, c->gas.gamma, c->stream);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->state_ok = true;
    return TY_OK;
}

int ty_state_get(ty_ctx* c, double* rho, double* rhoE, double* mom)
{
    if (!c || !rho || !rhoE || !mom) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    const int n = c->n, nv = c->nv;
    std::vector<double> h((size_t)nv * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_u0, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
        int o = c->perm[i];
        rho[o] = h[i];
        rhoE[o] = h[(size_t)n + i];
        for (int d = 0; d < 3; ++d) mom[(size_t)3 * o + d] = d < c->dim ? h[(size_t)(2 + d) * n + i] : 0.0;
    }
    return TY_OK;
}

namespace {

template <typename R>
KArgs<R> make_args(ty_ctx* c, const Geo<R>& g)
{
    KArgs<R> a;
    memset(&a, 0, sizeof(a));
    a.n = c->n;
    a.ns = c->ns;
    a.dense = c->dense;
    a.nbr = c->d_nbr;
    a.fw = c->d_fw;
    a.link = c->d_link;
    a.bnd_bits = c->d_bnd_bits;
    a.g = g;
    a.boco = c->d_boco;
    a.vol = c->d_vol;
    a.ph = c->ph;
    a.us = (const R*)c->d_us[c->cur];
    a.us_next = (R*)c->d_us[c->cur ^ 1];
    a.u0 = c->d_u0;
    a.acc = c->d_acc;
    // TY_PREC_MIXED_F32SUM: the FP32 kernels keep the sum in the first half of the same buffer
    a.acc32 = sizeof(R) == sizeof(float) && c->acc32 ? (float*)c->d_acc : nullptr;
    a.rstore = c->d_rstore;
    const int prec = sizeof(R) == sizeof(float) ? 1 : 0;
    a.grad = (const R*)c->work_grad[prec];
    a.grad_out = (R*)c->work_grad[prec];
    // residual evaluations of a viscous scheme pass the viscous records from K1 to K2
    a.vrec = c->ph.viscous ? 1 : 0;
    a.gbnd = (const R*)c->work_gbnd[prec];
    a.gbnd_out = (R*)c->work_gbnd[prec];
    a.qf = (const R*)c->work_qf[prec];
    a.qf_out = (R*)c->work_qf[prec];  // null for first-order schemes
    a.qf_mono = (R*)c->work_qf2[prec];  // null unless MONOTONIC1
    a.res_out = c->d_res;
    a.d_dt = c->d_dt;
    if (c->steady_on) {
        a.dtloc = c->steady_local ? c->d_dtloc : nullptr;
        a.rnorm = c->d_rnorm;
        a.stop = &c->d_steady->stop;
        for (int k = 0; k < 3; ++k) a.rref[k] = c->steady_ref[k];
    }
    a.store_slot = -1;
    for (int j = 0; j < MAX_STAGES; ++j) a.read_slot[j] = -1;
    return a;
}

// a timing event pair: both created and the first recorded on the context stream, or neither kept
hipError_t ev_begin(ty_ctx* c, hipEvent_t* e0, hipEvent_t* e1)
{
    *e0 = *e1 = nullptr;
    hipError_t e = hipEventCreate(e0);
    if (e == hipSuccess) e = hipEventCreate(e1);
    if (e == hipSuccess) e = hipEventRecord(*e0, c->stream);
    if (e != hipSuccess) {
        if (*e0) (void)hipEventDestroy(*e0);
        if (*e1) (void)hipEventDestroy(*e1);
        *e0 = *e1 = nullptr;
    }
    return e;
}

// records the second event and hands the pair to the list (ty_timing or free_state destroys it)
hipError_t ev_end(ty_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>& ev, hipEvent_t e0, hipEvent_t e1)
{
    hipError_t e = hipEventRecord(e1, c->stream);
    ev.push_back({e0, e1});
    return e;
}

template <typename R>
hipError_t timed(ty_ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>& ev, hipError_t (*fn)(const KArgs<R>&, hipStream_t),
                 const KArgs<R>& a)
{
    if (!c->timing) return fn(a, c->stream);
    hipEvent_t e0, e1;
    hipError_t e = ev_begin(c, &e0, &e1);
    if (e != hipSuccess) return e;
    e = fn(a, c->stream);
    const hipError_t er = ev_end(c, ev, e0, e1);
    return e != hipSuccess ? e : er;
}

// residual kernels of one evaluation on the current stage state
template <int DIM, typename R>
hipError_t eval_kernels(ty_ctx* c, KArgs<R>& a)
{
    hipError_t e = hipSuccess;
    const bool grad = c->ph.method != 'N' || c->ph.viscous;
    if (grad) {
        e = timed<R>(c, c->ev_grad, &launch_gradient<DIM, R>, a);
        if (e != hipSuccess) return e;
    }
    return timed<R>(c, c->ev_flux, &launch_flux<DIM, R>, a);
}

template <int DIM, typename R>
hipError_t rk_stages(ty_ctx* c, const Geo<R>& g)
{
    const int ns = c->sch.nstage;
    auto coef = [&](int i, int j) { return c->rk[(size_t)(i - 1) * ns + (j - 1)]; };
    // storage slots
    int slot_of[MAX_STAGES + 1];
    int nslot = 0;
    for (int s = 1; s <= ns; ++s) {
        slot_of[s] = -1;
        for (int t = s + 1; t < ns; ++t)
            if (coef(t, s) != 0.0) { slot_of[s] = nslot++; break; }
    }
    for (int s = 1; s <= ns; ++s) {
        KArgs<R> a = make_args<R>(c, g);
        a.mode = 0;
        a.stage = s;
        a.nstage = ns;
        a.b = coef(ns, s);
        a.store_slot = slot_of[s];
        if (s < ns) {
            for (int j = 1; j <= s; ++j) a.a[j - 1] = coef(s, j);
            for (int j = 1; j < s; ++j) a.read_slot[j - 1] = coef(s, j) != 0.0 ? slot_of[j] : -1;
        }
        hipError_t e = eval_kernels<DIM, R>(c, a);
        if (e != hipSuccess) return e;
        c->cur ^= 1;
    }
    return hipSuccess;
}

int run_step(ty_ctx* c)
{
    hipError_t e;
    if (c->mixed) e = c->dim == 2 ? rk_stages<2, float>(c, c->geo_f) : rk_stages<3, float>(c, c->geo_f);
    else e = c->dim == 2 ? rk_stages<2, double>(c, c->geo_d) : rk_stages<3, double>(c, c->geo_d);
    if (e != hipSuccess) return fail(c, TY_ERR_HIP, std::string("RK stage launch: ") + hipGetErrorString(e));
    return TY_OK;
}

int launch_dt(ty_ctx* c, double dtmax)
{
    DtArgs d;
    memset(&d, 0, sizeof(d));
    d.n = c->n;
    d.u0 = c->d_u0;
    d.vol = c->d_vol;
    d.sumA = c->d_sumA;
    d.cfl = c->d_cfl;
    d.gamma = c->gas.gamma;
    d.part = c->d_part;
    d.bad = c->d_bad;
    if (c->steady_on) {
        d.cfl = &c->d_steady->cfl;
        d.dtloc = c->steady_local ? c->d_dtloc : nullptr;
        d.stop = &c->d_steady->stop;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) HIPCHK(c, ev_begin(c, &e0, &e1));
    hipError_t e = c->dim == 2 ? launch_timestep<2>(d, dtmax, c->d_dt, c->d_err, c->stream)
                               : launch_timestep<3>(d, dtmax, c->d_dt, c->d_err, c->stream);
    if (c->timing) {
        const hipError_t er = ev_end(c, c->ev_dt, e0, e1);
        if (e == hipSuccess) e = er;
    }
    if (e != hipSuccess) return fail(c, TY_ERR_HIP, std::string("time step launch: ") + hipGetErrorString(e));
    return TY_OK;
}

int check_err(ty_ctx* c)
{
    int h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h) {
        HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return fail(c, TY_ERR_STATE, "negative or NaN internal energy (calc_ns_timestep)");
    }
    return TY_OK;
}

}  // namespace

int ty_timestep(ty_ctx* c, double dtmax, double* dt)
{
    if (!c || !dt) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = launch_dt(c, dtmax);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(&c->h_dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    rc = check_err(c);
    if (rc) return rc;
    *dt = c->h_dt;
    return TY_OK;
}

int ty_residual(ty_ctx* c, double dt, double* res_rho, double* res_rhoE, double* res_mom)
{
    if (!c || !res_rho || !res_rhoE || !res_mom) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipError_t e;
    if (c->mixed) {
        KArgs<float> a = make_args<float>(c, c->geo_f);
        a.mode = 1;
        e = c->dim == 2 ? eval_kernels<2, float>(c, a) : eval_kernels<3, float>(c, a);
    } else {
        KArgs<double> a = make_args<double>(c, c->geo_d);
        a.mode = 1;
        e = c->dim == 2 ? eval_kernels<2, double>(c, a) : eval_kernels<3, double>(c, a);
    }
    HIPCHK(c, e);
    const int n = c->n, nv = c->nv;
    std::vector<double> h((size_t)nv * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_res, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
        int o = c->perm[i];
        res_rho[o] = h[i];
        res_rhoE[o] = h[(size_t)n + i];
        for (int d = 0; d < 3; ++d) res_mom[(size_t)3 * o + d] = d < c->dim ? h[(size_t)(2 + d) * n + i] : 0.0;
    }
    return TY_OK;
}

int ty_rk_step(ty_ctx* c, double dt)
{
    if (!c) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = run_step(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TY_OK;
}

int ty_advance(ty_ctx* c, int32_t nsteps, double dtmax, double* dt_last)
{
    if (!c || nsteps < 0) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    for (int s = 0; s < nsteps; ++s) {
        int rc = launch_dt(c, dtmax);
        if (rc) return rc;
        rc = run_step(c);
        if (rc) return rc;
    }
    if (dt_last) {
        HIPCHK(c, hipMemcpyAsync(&c->h_dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        int rc = check_err(c);
        if (rc) return rc;
        *dt_last = c->h_dt;
    }
    return TY_OK;
}

int ty_sync(ty_ctx* c)
{
    if (!c) return TY_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_err) return check_err(c);
    return TY_OK;
}

// ---------------------------------------------------------------------------
// steady mode: reference values, local time steps, the device-resident steady cycle

int ty_refcons(ty_ctx* c, double* out)
{
    if (!c || !out) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    const int n = c->n;
    std::vector<double> h((size_t)c->nv * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_u0, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // calc_refcons.f90:46-105 in the caller's cell order, define_refcons (MENU_SOLVER.f90:262-276)
    double v = 0.0, s0 = 0.0, s1 = 0.0;
    for (int o = 0; o < n; ++o) {
        const int i = c->iperm[o];
        v += c->vol[i];
        s0 += h[i] * c->vol[i];
        s1 += h[(size_t)n + i] * c->vol[i];
    }
    out[0] = s0 / v;
    out[1] = s1 / v;
    out[2] = std::sqrt(out[0] * out[1]);
    return TY_OK;
}

namespace {

int ensure_steady(ty_ctx* c, int maxit)
{
    int rc;
    if (!c->d_dtloc) {
        if ((rc = dalloc(c, &c->d_dtloc, c->n)) || (rc = dalloc(c, &c->d_rnorm, c->n)) ||
            (rc = dalloc(c, &c->d_spart, (c->n + BLOCK - 1) / BLOCK)) || (rc = dalloc(c, &c->d_steady, 1)))
            return rc;
    }
    if (c->hist_len < maxit) {
        if ((rc = dalloc(c, &c->d_hist, maxit))) return rc;
        c->hist_len = maxit;
    }
    return TY_OK;
}

}  // namespace

int ty_local_timestep(ty_ctx* c, double* dt_cells, double* dt_min)
{
    if (!c || !dt_cells) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_steady(c, 1))) return rc;
    // the time-step kernels with per-cell output (loc_stab_cond) at the scheme's CFL number
    DtArgs d;
    memset(&d, 0, sizeof(d));
    d.n = c->n;
    d.u0 = c->d_u0;
    d.vol = c->d_vol;
    d.sumA = c->d_sumA;
    d.cfl = c->d_cfl;
    d.gamma = c->gas.gamma;
    d.part = c->d_part;
    d.bad = c->d_bad;
    d.dtloc = c->d_dtloc;
    const hipError_t e = c->dim == 2 ? launch_timestep<2>(d, HUGE_VAL, c->d_dt, c->d_err, c->stream)
                                     : launch_timestep<3>(d, HUGE_VAL, c->d_dt, c->d_err, c->stream);
    HIPCHK(c, e);
    std::vector<double> h(c->n);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_dtloc, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->h_dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((rc = check_err(c))) return rc;
    for (int i = 0; i < c->n; ++i) dt_cells[c->perm[i]] = h[i];
    if (dt_min) *dt_min = c->h_dt;
    return TY_OK;
}

int ty_steady_cycle(ty_ctx* c, const ty_steady* st, double* res_hist, ty_steady_stats* stats)
{
    if (!c || !st) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    if (st->maxit < 1 || st->check_every < 1 || !(st->cfl > 0.0) || !(st->cfl_max > 0.0) || !(st->residu_ref > 0.0) ||
        !(st->refsca[0] > 0.0) || !(st->refsca[1] > 0.0) || !(st->refvec > 0.0))
        return fail(c, TY_ERR_ARG, "steady cycle: bad parameters");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_steady(c, st->maxit))) return rc;
    // integration_cyclezone.f90:50: the cycle starts with cur_res = residu_ref, so the first
    // iteration's CFL ramp term log10(cur_res / residu_ref) is zero
    SteadyDev h0;
    h0.cur_res = h0.residu_ref = st->residu_ref;
    h0.cfl = std::min(st->cfl * std::max(1.0, 1.0 - st->wlog), st->cfl_max);
    h0.iter = 0;
    h0.stop = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_steady, &h0, sizeof(h0), hipMemcpyHostToDevice, c->stream));
    SteadyArgs sa;
    sa.stabnb = st->cfl;
    sa.cfl_max = st->cfl_max;
    sa.maxres = st->maxres;
    sa.wlog = st->wlog;
    sa.maxit = st->maxit;
    sa.nblk = (c->n + BLOCK - 1) / BLOCK;
    sa.st = c->d_steady;
    sa.part = c->d_spart;
    sa.hist = c->d_hist;
    c->steady_on = true;
    c->steady_local = st->local_dt != 0;
    c->steady_ref[0] = st->refsca[0];
    c->steady_ref[1] = st->refsca[1];
    c->steady_ref[2] = st->refvec;
    // iterations are queued without host synchronisation; once the device sets the stop flag
    // (convergence or maxit), the later launches do nothing; the host looks every check_every
    for (int it = 0; it < st->maxit && rc == TY_OK; ++it) {
        if ((rc = launch_dt(c, HUGE_VAL)) || (rc = run_step(c))) break;
        const hipError_t e = launch_steady_update(c->d_rnorm, c->n, sa, c->d_spart, c->stream);
        if (e != hipSuccess) {
            rc = fail(c, TY_ERR_HIP, std::string("steady update launch: ") + hipGetErrorString(e));
            break;
        }
        if ((it + 1) % st->check_every == 0 && it + 1 < st->maxit) {
            int stop = 0;
            const hipError_t e2 = hipMemcpyAsync(&stop, &c->d_steady->stop, sizeof(int), hipMemcpyDeviceToHost, c->stream);
            if (e2 != hipSuccess) {
                rc = fail(c, TY_ERR_HIP, std::string("steady stop flag: ") + hipGetErrorString(e2));
                break;
            }
            if ((rc = check_err(c))) break;  // synchronises; a bad state stops the cycle
            if (stop) break;
        }
    }
    c->steady_on = false;
    c->steady_local = false;
    if (rc) return rc;
    // stages after the stop did nothing: the stage records follow the FP64 state again
    c->cur = 0;
    hipError_t e = c->mixed ? launch_to_stage<float>(c->d_u0, (float*)c->d_us[0], c->nv, c->n, c->gas.gamma, c->stream)
                            : launch_to_stage<double>(c->d_u0, (double*)c->d_us[0], c->nv, c->n, c->gas.gamma, c->stream);
    HIPCHK(c, e);
    SteadyDev h1;
    HIPCHK(c, hipMemcpyAsync(&h1, c->d_steady, sizeof(h1), hipMemcpyDeviceToHost, c->stream));
    if ((rc = check_err(c))) return rc;
    if (res_hist && h1.iter > 0)
        HIPCHK(c, hipMemcpy(res_hist, c->d_hist, (size_t)h1.iter * sizeof(double), hipMemcpyDeviceToHost));
    if (stats) {
        stats->iterations = h1.iter;
        stats->converged = h1.cur_res / h1.residu_ref <= st->maxres ? 1 : 0;
        stats->cur_res = h1.cur_res;
        stats->residu_ref = h1.residu_ref;
        stats->cfl = h1.cfl;
    }
    return TY_OK;
}

// ---------------------------------------------------------------------------
// volume probes (def_probe.f90:34-80, prb_zone_vol.f90:34-51)

int ty_probe_define(ty_ctx* c, const char* quantity, int32_t type, int32_t* id)
{
    if (!c || !quantity || !id) return TY_ERR_ARG;
    if (type != PROBE_AVG && type != PROBE_MIN && type != PROBE_MAX)
        return fail(c, TY_ERR_ARG, "probe type must be VOL_AVERAGE (21), VOL_MIN (22) or VOL_MAX (23)");
    ProbeProg p;
    std::string err;
    if (probe_compile(quantity, type, &p, &err)) return fail(c, TY_ERR_ARG, "probe \"" + std::string(quantity) + "\": " + err);
    c->probes.push_back(p);
    *id = (int32_t)c->probes.size() - 1;
    return TY_OK;
}

int ty_probe_eval(ty_ctx* c, double t, int32_t nprobe, const int32_t* ids, double* results)
{
    if (!c || nprobe < 0 || (nprobe > 0 && (!ids || !results))) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    for (int i = 0; i < nprobe; ++i)
        if (ids[i] < 0 || ids[i] >= (int)c->probes.size()) return fail(c, TY_ERR_ARG, "unknown probe id");
    if (nprobe == 0) return TY_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const int nb = probe_blocks(c->n);
    if (c->probe_cap < nprobe) {
        int rc;
        if ((rc = dalloc(c, &c->d_probe, (size_t)2 * nb + nprobe))) return rc;
        c->probe_cap = nprobe;
    }
    ProbeData d;
    d.n = c->n;
    d.dim = c->dim;
    d.u0 = c->d_u0;
    d.cc = c->d_cc;
    d.vol = c->d_vol;
    d.g1 = c->gas.gamma - 1.0;
    d.t = t;
    double* out = c->d_probe + (size_t)2 * nb;
    for (int i = 0; i < nprobe; ++i) HIPCHK(c, launch_probe(c->probes[ids[i]], d, c->d_probe, out + i, c->stream));
    HIPCHK(c, hipMemcpyAsync(results, out, (size_t)nprobe * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TY_OK;
}

int ty_probe_eval_noalloc(ty_ctx* c, double t, int32_t nprobe, const int32_t* ids, double* results)
{
    if (!c || nprobe < 0 || (nprobe > 0 && (!ids || !results))) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    if (c->n > PROBE_SCRATCH_MAX_CELLS)
        return fail(c, TY_ERR_ARG, "ty_probe_eval_noalloc: " + std::to_string(c->n) + " cells exceed its workspace limit of " +
                                       std::to_string(PROBE_SCRATCH_MAX_CELLS));
    if (nprobe > PROBE_SCRATCH_MAX_PROBES)
        return fail(c, TY_ERR_ARG, "ty_probe_eval_noalloc: " + std::to_string(nprobe) + " probes exceed its limit of " +
                                       std::to_string(PROBE_SCRATCH_MAX_PROBES) + " per call");
    for (int i = 0; i < nprobe; ++i)
        if (ids[i] < 0 || ids[i] >= (int)c->probes.size()) return fail(c, TY_ERR_ARG, "unknown probe id");
    if (nprobe == 0) return TY_OK;
    HIPCHK(c, hipSetDevice(c->device));
    ProbeData d;
    d.n = c->n;
    d.dim = c->dim;
    d.u0 = c->d_u0;
    d.cc = c->d_cc;
    d.vol = c->d_vol;
    d.g1 = c->gas.gamma - 1.0;
    d.t = t;
    // the same kernels as ty_probe_eval with the partials and results in the fixed workspace
    double* part = c->d_scratch;
    double* out = c->d_scratch + (size_t)2 * (PROBE_SCRATCH_MAX_CELLS / 256);
    for (int i = 0; i < nprobe; ++i) HIPCHK(c, launch_probe(c->probes[ids[i]], d, part, out + i, c->stream));
    HIPCHK(c, hipMemcpyAsync(results, out, (size_t)nprobe * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TY_OK;
}

// Per-kernel timing for the benchmark harness: when enabled, every gradient,
// flux and time-step launch is bracketed by HIP events on the context stream.
int ty_timing(ty_ctx* c, int32_t enable, double* ms, int32_t* count)
{
    if (!c) return TY_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<std::pair<hipEvent_t, hipEvent_t>>* lists[5] = {&c->ev_grad, &c->ev_flux, &c->ev_dt, &c->ev_tan,
                                                                &c->ev_adj};
    for (int k = 0; k < 5; ++k) {
        double tot = 0.0;
        hipError_t err = hipSuccess;
        for (auto& e : *lists[k]) {
            float t = 0.f;
            const hipError_t et = hipEventElapsedTime(&t, e.first, e.second);
            if (et == hipSuccess) tot += t;
            else if (err == hipSuccess) err = et;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (err != hipSuccess) {
            lists[k]->clear();
            return fail(c, TY_ERR_HIP, std::string("hipEventElapsedTime: ") + hipGetErrorString(err));
        }
        if (ms) ms[k] = tot;
        if (count) count[k] = (int32_t)lists[k]->size();
        lists[k]->clear();
    }
    c->timing = enable != 0;
    return TY_OK;
}

// ---------------------------------------------------------------------------
// Matrix-free implicit solve: (I - J) x = b, J = dR/dU, R = -dt/V sum F.A.
// FP64 iterative refinement (central-difference JVP of the FP64 residual)
// around restarted, right-preconditioned GMRES whose Krylov vectors, block-Jacobi
// preconditioner and one-sided JVP run in the inner precision.  Classical
// Gram-Schmidt is applied twice per Arnoldi step (two fused multi-dot/update
// passes) instead of the reference's modified Gram-Schmidt (bdlu_gmres.f90:101-107).

namespace {

template <typename R>
const Geo<R>& geo_of(ty_ctx* c);
template <>
const Geo<double>& geo_of<double>(ty_ctx* c) { return c->geo_d; }
template <>
const Geo<float>& geo_of<float>(ty_ctx* c) { return c->geo_f; }

template <typename R>
hipError_t eval_residual(ty_ctx* c, const R* state, double* out, R* dinv, double* diag)
{
    KArgs<R> a = make_args<R>(c, geo_of<R>(c));
    a.us = state;
    a.mode = 1;
    a.res_out = out;
    a.dinv_out = dinv;
    a.diag_out = diag;
    return c->dim == 2 ? eval_kernels<2, R>(c, a) : eval_kernels<3, R>(c, a);
}

int alloc_krylov(ty_ctx* c, int m, int prec)
{
    if (c->kry_ok && c->kry_m >= m && c->kry_prec == prec) return TY_OK;
    int rc;
    if ((rc = ensure_work(c, prec)) || (rc = ensure_work(c, 0))) return rc;
    const long len = (long)c->nv * c->n;
    const size_t pb = prec ? sizeof(float) : sizeof(double);
    if (!c->kry_base) {
        if ((rc = dalloc(c, &c->k_b, len)) || (rc = dalloc(c, &c->k_x, len)) || (rc = dalloc(c, &c->k_r, len)) ||
            (rc = dalloc(c, &c->k_r0, len)) || (rc = dalloc(c, &c->k_r1, len)) || (rc = dalloc(c, &c->k_r2, len)) ||
            (rc = dalloc(c, &c->k_pd, len)) || (rc = dalloc(c, &c->k_dx, len)) || (rc = dalloc(c, &c->k_pf, len)) ||
            (rc = dalloc(c, &c->k_part, (size_t)krylov_partials(len))) || (rc = dalloc(c, &c->k_h, 3 * MAXK + 4)))
            return rc;
        c->kry_base = true;
    }
    if (c->kry_ok) {
        // a larger basis or another inner precision: release the previous basis and work vectors
        for (void** q : {&c->k_V, &c->k_z, &c->k_w, &c->k_t, &c->k_dinv}) {
            auto it = std::find(c->owned.begin(), c->owned.end(), *q);
            if (it != c->owned.end()) c->owned.erase(it);
            HIPCHK(c, hipFree(*q));
            *q = nullptr;
        }
        c->kry_ok = false;
    }
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, (size_t)(m + 1) * len * pb));
    c->owned.push_back(p);
    c->k_V = p;
    for (void** q : {&c->k_z, &c->k_w, &c->k_t}) {
        HIPCHK(c, hipMalloc(&p, (size_t)len * pb));
        c->owned.push_back(p);
        *q = p;
    }
    HIPCHK(c, hipMalloc(&p, (size_t)c->nv * c->nv * c->n * pb));
    c->owned.push_back(p);
    c->k_dinv = p;
    c->kry_ok = true;
    c->kry_m = m;
    c->kry_prec = prec;
    return TY_OK;
}

template <typename T>
int dot_host(ty_ctx* c, const T* a, const T* b, long len, double* out)
{
    HIPCHK(c, multidot<T>(a, b, 0, 1, len, c->k_part, c->k_h, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->k_h, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TY_OK;
}

template <typename P>
P* pert_buffer(ty_ctx* c);
template <>
double* pert_buffer<double>(ty_ctx* c) { return c->k_pd; }
template <>
float* pert_buffer<float>(ty_ctx* c) { return c->k_pf; }

// restarted GMRES in precision P on (I - J) d = rhs; d accumulates into c->k_dx (FP64)
template <typename P>
int gmres(ty_ctx* c, const double* rhs, const ty_krylov* k, double unorm, int* iters, double* rel)
{
    const long len = (long)c->nv * c->n;
    const int m = k->nkrylov;
    P* V = (P*)c->k_V;
    P* z = (P*)c->k_z;
    P* w = (P*)c->k_w;
    P* t = (P*)c->k_t;
    const P* dinv = (const P*)c->k_dinv;
    P* pert = pert_buffer<P>(c);
    const double sq = std::sqrt(sizeof(P) == sizeof(float) ? 1.1920928955078125e-07 : 2.220446049250313e-16);
    int rc;
    double bb;
    if ((rc = dot_host<double>(c, rhs, rhs, len, &bb))) return rc;
    const double beta0 = std::sqrt(bb);
    *rel = 0.0;
    if (beta0 == 0.0) return TY_OK;
    const double target = k->inner_tol * beta0;
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m), y(m), hh(MAXK), h2(MAXK), hbuf(3 * MAXK);
    double res = beta0;
    HIPCHK(c, hipMemsetAsync(c->k_dx, 0, len * sizeof(double), c->stream));
    // wout = zin - J zin; the difference step eps = sqrt(machine eps) (1 + ||u||) / ||zin|| (RMS norms)
    // is formed on the device from the device norm, so the step needs no host round trip; for
    // zin = 0 both eps and 1/eps are 0 and wout = zin
    double* d_zz = c->k_h + 3 * MAXK;
    double* d_eps = d_zz + 1;  // eps, 1 / eps
    auto A_apply = [&](const P* zin, P* wout) -> int {
        HIPCHK(c, multidot<P>(zin, zin, 0, 1, len, c->k_part, d_zz, c->stream));
        HIPCHK(c, jvp_step(d_zz, len, sq * (1.0 + unorm), d_eps, c->stream));
        HIPCHK(c, (fill_state<P, P>(pert, c->d_u0, 0.0, zin, c->nv, c->n, c->gas.gamma, c->stream, d_eps)));
        HIPCHK(c, eval_residual<P>(c, pert, c->k_r1, nullptr, nullptr));
        HIPCHK(c, apply_a<P>(wout, zin, c->k_r1, c->k_r0, 0.0, len, c->stream, d_eps + 1));
        return TY_OK;
    };
    for (int cycle = 0; cycle < std::max(1, k->max_restart); ++cycle) {
        double beta;
        if (cycle == 0) {
            beta = beta0;
            HIPCHK(c, (scale<P, double>(V, rhs, 1.0 / beta, len, c->stream)));
        } else {
            // current residual rhs - (d - J d) from a one-sided JVP of the accumulated correction
            double dd;
            if ((rc = dot_host<double>(c, c->k_dx, c->k_dx, len, &dd))) return rc;
            const double dn = std::sqrt(dd / (double)len);
            const double eps = sq * (1.0 + unorm) / std::max(dn, 1e-300);
            HIPCHK(c, (fill_state<P, double>(pert, c->d_u0, eps, c->k_dx, c->nv, c->n, c->gas.gamma, c->stream)));
            HIPCHK(c, eval_residual<P>(c, pert, c->k_r1, nullptr, nullptr));
            HIPCHK(c, outer_residual(c->k_r2, rhs, c->k_dx, c->k_r1, c->k_r0, 1.0 / eps, len, c->stream));
            if ((rc = dot_host<double>(c, c->k_r2, c->k_r2, len, &bb))) return rc;
            beta = std::sqrt(bb);
            if (beta <= target) { res = beta; break; }
            HIPCHK(c, (scale<P, double>(V, c->k_r2, 1.0 / beta, len, c->stream)));
        }
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jj = 0;
        for (int j = 0; j < m; ++j) {
            P* vj = V + (size_t)j * len;
            const P* zj = vj;
            if (k->precond) {
                HIPCHK(c, block_apply<P>(c->nv, dinv, vj, z, c->n, c->stream));
                zj = z;
            }
            if ((rc = A_apply(zj, w))) return rc;
            // classical Gram-Schmidt, twice, in three sweeps over the basis: h1 = V^T w; w -= V h1
            // with h2 = V^T w in the same sweep (the dots a separate multi-dot would give, bit for
            // bit); w -= V h2 with ||w||^2 in the same sweep; one host round trip per step
            double* dh1 = c->k_h;
            double* dh2 = c->k_h + MAXK;
            double* dnn = c->k_h + 2 * MAXK;
            HIPCHK(c, multidot<P>(w, V, len, j + 1, len, c->k_part, dh1, c->stream));
            HIPCHK(c, cgs_pass<P>(w, V, len, dh1, j + 1, true, len, c->k_part, dh2, c->stream));
            HIPCHK(c, cgs_pass<P>(w, V, len, dh2, j + 1, false, len, c->k_part, dnn, c->stream));
            HIPCHK(c, hipMemcpyAsync(hbuf.data(), c->k_h, hbuf.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int i = 0; i <= j; ++i) {
                hh[i] = hbuf[i];
                h2[i] = hbuf[MAXK + i];
            }
            const double ww = hbuf[2 * MAXK];
            const double hn = std::sqrt(ww);
            double* col = &H[(size_t)j * (m + 1)];
            for (int i = 0; i <= j; ++i) col[i] = hh[i] + h2[i];
            col[j + 1] = hn;
            for (int i = 0; i < j; ++i) {
                const double tmp = col[i];
                col[i] = cs[i] * tmp + sn[i] * col[i + 1];
                col[i + 1] = -sn[i] * tmp + cs[i] * col[i + 1];
            }
            const double rr = std::sqrt(col[j] * col[j] + col[j + 1] * col[j + 1]);
            cs[j] = rr > 0 ? col[j] / rr : 1.0;
            sn[j] = rr > 0 ? col[j + 1] / rr : 0.0;
            col[j] = rr;
            col[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            res = std::fabs(g[j + 1]);
            ++*iters;
            jj = j + 1;
            if (res <= target || hn == 0.0) break;
            if (j + 1 < m) HIPCHK(c, (scale<P, P>(V + (size_t)(j + 1) * len, w, 1.0 / hn, len, c->stream)));
        }
        // y = H^-1 g, correction d += M^-1 V y
        for (int i = jj - 1; i >= 0; --i) {
            double tmp = g[i];
            for (int l = i + 1; l < jj; ++l) tmp -= H[(size_t)l * (m + 1) + i] * y[l];
            y[i] = tmp / H[(size_t)i * (m + 1) + i];
        }
        HIPCHK(c, hipMemcpyAsync(c->k_h, y.data(), jj * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, lincomb<P>(t, V, len, c->k_h, jj, len, c->stream));
        if (k->precond) {
            HIPCHK(c, block_apply<P>(c->nv, dinv, t, z, c->n, c->stream));
            HIPCHK(c, accumulate<P>(c->k_dx, z, len, c->stream));
        } else {
            HIPCHK(c, accumulate<P>(c->k_dx, t, len, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (res <= target) break;
    }
    *rel = res / beta0;
    return TY_OK;
}

// solve (I - J) x = k_b at the current state into k_x
int solve_dev(ty_ctx* c, double dt, const ty_krylov* k, ty_krylov_stats* st)
{
    if (k->nkrylov < 1 || k->nkrylov >= MAXK) return fail(c, TY_ERR_ARG, "nkrylov must be in 1..31");
    const int prec = k->inner_precision == TY_PREC_MIXED ? 1 : 0;
    int rc;
    if ((rc = alloc_krylov(c, k->nkrylov, prec))) return rc;
    const long len = (long)c->nv * c->n;
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    double uu, bb;
    if ((rc = dot_host<double>(c, c->d_u0, c->d_u0, len, &uu)) || (rc = dot_host<double>(c, c->k_b, c->k_b, len, &bb)))
        return rc;
    const double unorm = std::sqrt(uu / (double)len);
    const double bnorm = std::sqrt(bb);
    // inner-precision residual at U (base of the one-sided JVP) and the preconditioner
    if (prec) {
        HIPCHK(c, (fill_state<float, double>(c->k_pf, c->d_u0, 0.0, c->d_u0, c->nv, c->n, c->gas.gamma, c->stream)));
        HIPCHK(c, eval_residual<float>(c, c->k_pf, c->k_r0, k->precond ? (float*)c->k_dinv : nullptr, nullptr));
    } else {
        HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, 0.0, c->d_u0, c->nv, c->n, c->gas.gamma, c->stream)));
        HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_r0, k->precond ? (double*)c->k_dinv : nullptr, nullptr));
    }
    HIPCHK(c, hipMemsetAsync(c->k_x, 0, len * sizeof(double), c->stream));
    HIPCHK(c, hipMemcpyAsync(c->k_r, c->k_b, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    st->outer_iters = 0;
    st->inner_iters = 0;
    st->rhs_norm = bnorm;
    st->rel_res = bnorm > 0 ? 1.0 : 0.0;
    if (bnorm == 0.0) return TY_OK;
    const double cb = std::cbrt(2.220446049250313e-16);
    for (int outer = 0;; ++outer) {
        if (outer > 0) {
            // FP64 residual of the refinement with a central-difference JVP
            double xx;
            if ((rc = dot_host<double>(c, c->k_x, c->k_x, len, &xx))) return rc;
            const double xn = std::sqrt(xx / (double)len);
            const double eps = cb * (1.0 + unorm) / std::max(xn, 1e-300);
            HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, eps, c->k_x, c->nv, c->n, c->gas.gamma, c->stream)));
            HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_r1, nullptr, nullptr));
            HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, -eps, c->k_x, c->nv, c->n, c->gas.gamma, c->stream)));
            HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_r2, nullptr, nullptr));
            HIPCHK(c, outer_residual(c->k_r, c->k_b, c->k_x, c->k_r1, c->k_r2, 0.5 / eps, len, c->stream));
        }
        double rr;
        if ((rc = dot_host<double>(c, c->k_r, c->k_r, len, &rr))) return rc;
        st->rel_res = std::sqrt(rr) / bnorm;
        if (st->rel_res <= k->outer_tol || outer >= k->max_outer) break;
        int it = 0;
        double irel = 0.0;
        if (prec) {
            rc = gmres<float>(c, c->k_r, k, unorm, &it, &irel);
        } else {
            rc = gmres<double>(c, c->k_r, k, unorm, &it, &irel);
        }
        if (rc) return rc;
        st->inner_iters += it;
        st->outer_iters = outer + 1;
        HIPCHK(c, accumulate<double>(c->k_x, c->k_dx, len, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return st->rel_res <= k->outer_tol ? TY_OK : TY_ERR_NOT_CONVERGED;
}

void host_to_dev_vec(ty_ctx* c, const double* h, std::vector<double>& d)
{
    const int n = c->n;
    d.assign((size_t)c->nv * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int o = c->perm[i];
        for (int v = 0; v < c->nv; ++v) d[(size_t)v * n + i] = h[(size_t)v * n + o];
    }
}

void dev_to_host_vec(ty_ctx* c, const std::vector<double>& d, double* h)
{
    const int n = c->n;
    for (int i = 0; i < n; ++i) {
        const int o = c->perm[i];
        for (int v = 0; v < 5; ++v) h[(size_t)v * n + o] = v < c->nv ? d[(size_t)v * n + i] : 0.0;
    }
}

}  // namespace

int ty_jacobian_diag(ty_ctx* c, double dt, double* diag)
{
    if (!c || !diag) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    const int n = c->n, nv = c->nv;
    double* d_diag = nullptr;
    double* d_inv = nullptr;
    double* d_st = nullptr;
    int rc;
    if ((rc = dalloc(c, &d_diag, (size_t)nv * nv * n)) || (rc = dalloc(c, &d_inv, (size_t)nv * nv * n)) ||
        (rc = dalloc(c, &d_st, (size_t)nv * n)))
        return rc;
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = ensure_work(c, 0))) return rc;
    HIPCHK(c, launch_to_stage<double>(c->d_u0, d_st, nv, n, c->gas.gamma, c->stream));
    HIPCHK(c, eval_residual<double>(c, d_st, c->d_res, d_inv, d_diag));
    std::vector<double> h((size_t)nv * nv * n);
    HIPCHK(c, hipMemcpyAsync(h.data(), d_diag, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
        const int o = c->perm[i];
        double* out = diag + (size_t)25 * o;
        for (int k = 0; k < 25; ++k) out[k] = (k % 6 == 0) ? 1.0 : 0.0;
        for (int a = 0; a < nv; ++a)
            for (int b = 0; b < nv; ++b) out[a * 5 + b] = h[(size_t)(a * nv + b) * n + i];
    }
    // release the temporaries
    for (double* p : {d_diag, d_inv, d_st}) {
        (void)hipFree(p);
        c->owned.erase(std::find(c->owned.begin(), c->owned.end(), (void*)p));
    }
    return TY_OK;
}

int ty_jvp(ty_ctx* c, double dt, const double* v, double* jv)
{
    if (!c || !v || !jv) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = alloc_krylov(c, std::max(1, c->kry_m), c->kry_prec < 0 ? 0 : c->kry_prec))) return rc;
    const long len = (long)c->nv * c->n;
    std::vector<double> h;
    host_to_dev_vec(c, v, h);
    HIPCHK(c, hipMemcpyAsync(c->k_x, h.data(), len * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    double uu, xx;
    if ((rc = dot_host<double>(c, c->d_u0, c->d_u0, len, &uu)) || (rc = dot_host<double>(c, c->k_x, c->k_x, len, &xx)))
        return rc;
    const double unorm = std::sqrt(uu / (double)len), xn = std::sqrt(xx / (double)len);
    const double eps = std::cbrt(2.220446049250313e-16) * (1.0 + unorm) / std::max(xn, 1e-300);
    HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, eps, c->k_x, c->nv, c->n, c->gas.gamma, c->stream)));
    HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_r1, nullptr, nullptr));
    HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, -eps, c->k_x, c->nv, c->n, c->gas.gamma, c->stream)));
    HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_r2, nullptr, nullptr));
    // Jv = (r1 - r2) / (2 eps): reuse outer_residual with b = 0 and x = 0 -> r = (r1 - r2) s
    HIPCHK(c, hipMemsetAsync(c->k_b, 0, len * sizeof(double), c->stream));
    HIPCHK(c, hipMemsetAsync(c->k_dx, 0, len * sizeof(double), c->stream));
    HIPCHK(c, outer_residual(c->k_r, c->k_b, c->k_dx, c->k_r1, c->k_r2, 0.5 / eps, len, c->stream));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->k_r, len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_to_host_vec(c, h, jv);
    return TY_OK;
}

int ty_linear_solve(ty_ctx* c, double dt, const ty_krylov* k, const double* b, double* x, ty_krylov_stats* st)
{
    if (!c || !k || !b || !x || !st) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = alloc_krylov(c, k->nkrylov, k->inner_precision == TY_PREC_MIXED ? 1 : 0))) return rc;
    const long len = (long)c->nv * c->n;
    std::vector<double> h;
    host_to_dev_vec(c, b, h);
    HIPCHK(c, hipMemcpyAsync(c->k_b, h.data(), len * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const int rs = solve_dev(c, dt, k, st);
    if (rs && rs != TY_ERR_NOT_CONVERGED) return rs;
    HIPCHK(c, hipMemcpyAsync(h.data(), c->k_x, len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_to_host_vec(c, h, x);
    if (rs) return fail(c, rs, "implicit solve did not reach the requested tolerance");
    return TY_OK;
}

int ty_implicit_step(ty_ctx* c, double dt, const ty_krylov* k, ty_krylov_stats* st)
{
    if (!c || !k || !st) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = alloc_krylov(c, k->nkrylov, k->inner_precision == TY_PREC_MIXED ? 1 : 0))) return rc;
    const long len = (long)c->nv * c->n;
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    // b = R(U) in FP64 (calc_rhs.f90:119-145)
    HIPCHK(c, (fill_state<double, double>(c->k_pd, c->d_u0, 0.0, c->d_u0, c->nv, c->n, c->gas.gamma, c->stream)));
    HIPCHK(c, eval_residual<double>(c, c->k_pd, c->k_b, nullptr, nullptr));
    const int rs = solve_dev(c, dt, k, st);
    if (rs && rs != TY_ERR_NOT_CONVERGED) return rs;
    // update_field.f90:38: U += dU, then refresh the stage-state copy
    HIPCHK(c, accumulate<double>(c->d_u0, c->k_x, len, c->stream));
    hipError_t e = c->mixed ? launch_to_stage<float>(c->d_u0, (float*)c->d_us[c->cur], c->nv, c->n, c->gas.gamma, c->stream)
                            : launch_to_stage<double>(c->d_u0, (double*)c->d_us[c->cur], c->nv, c->n, c->gas.gamma, c->stream);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rs) return fail(c, rs, "implicit solve did not reach the requested tolerance");
    return TY_OK;
}

// ---------------------------------------------------------------------------
// tangent-linear and adjoint of the first-order residual

namespace {

// HIP events around a launch when timing is enabled
struct EventScope {
    ty_ctx* c;
    std::vector<std::pair<hipEvent_t, hipEvent_t>>& ev;
    hipEvent_t e0 = nullptr;
    hipEvent_t e1 = nullptr;
    // timing only: a pair that cannot be created or recorded is dropped (the call itself goes on)
    EventScope(ty_ctx* cc, std::vector<std::pair<hipEvent_t, hipEvent_t>>& l) : c(cc), ev(l)
    {
        if (c->timing) (void)ev_begin(c, &e0, &e1);
    }
    ~EventScope()
    {
        if (c->timing && e0) (void)ev_end(c, ev, e0, e1);
    }
};

// device temporaries released when the call returns
struct DevTmp {
    ty_ctx* c;
    std::vector<void*> p;
    explicit DevTmp(ty_ctx* cc) : c(cc) {}
    ~DevTmp()
    {
        for (void* q : p) {
            (void)hipFree(q);
            auto it = std::find(c->owned.begin(), c->owned.end(), q);
            if (it != c->owned.end()) c->owned.erase(it);
        }
    }
    template <typename T>
    int alloc(T** out, size_t count)
    {
        const int rc = dalloc(c, out, count);
        if (!rc) p.push_back((void*)*out);
        return rc;
    }
};

int ad_prepare(ty_ctx* c, double dt, AdArgs& t, DevTmp& tmp)
{
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    if (c->ph.method != 'N')
        return fail(c, TY_ERR_UNSUPPORTED, "tangent/adjoint residual: first-order schemes only");
    if (c->ext_bc)
        return fail(c, TY_ERR_UNSUPPORTED, "tangent/adjoint residual: wall and subsonic-inlet boundaries not provided");
    HIPCHK(c, hipSetDevice(c->device));
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    memset(&t, 0, sizeof(t));
    t.k = make_args<double>(c, c->geo_d);
    t.bcell = c->d_bcell;
    t.nbcell = (int)c->bcell.size();
    t.fslot = c->d_fslot;
    t.k.us = c->d_u0;
    t.nf = c->nf;
    t.nb = c->nb;
    t.fcell = c->d_fcell;
    t.nfi = c->nfi;
    // primitive records of the state (the HLLC face kernels and the gathers read them)
    int rc;
    double* d_st = nullptr;
    if ((rc = tmp.alloc(&d_st, (size_t)c->nv * c->n))) return rc;
    HIPCHK(c, launch_to_stage<double>(c->d_u0, d_st, c->nv, c->n, c->gas.gamma, c->stream));
    t.qrec = d_st;
    if (c->ph.viscous) {
        // the cell gradients of the current state, as the residual kernels compute them
        if ((rc = ensure_work(c, 0))) return rc;
        double* d_vg = nullptr;
        if ((rc = tmp.alloc(&d_vg, (size_t)(c->dim * c->dim + c->dim) * c->n))) return rc;
        KArgs<double> a = make_args<double>(c, c->geo_d);
        a.us = d_st;
        a.qf_out = nullptr;
        a.vrec = 0;  // the derivative kernels read the full gradients of (rho, p, u..)
        a.vrec_out = d_vg;  // and the viscous records
        const hipError_t e = c->dim == 2 ? launch_gradient<2, double>(a, c->stream) : launch_gradient<3, double>(a, c->stream);
        HIPCHK(c, e);
        t.k.grad = (const double*)c->work_grad[0];
        t.vgrad = d_vg;
    }
    return TY_OK;
}

// work buffers of the residual adjoint (ty_residual_adjoint and each reversed stage of
// ty_unsteady_adjoint): per face side records, lambda dt/V records, (viscous) per face geometry,
// boundary-face gradient adjoints and the per-slot pushback
int alloc_adjoint_work(ty_ctx* c, AdArgs& t, DevTmp& tmp)
{
    const int n = c->n, nv = c->nv, dim = c->dim;
    int rc;
    double *d_face = nullptr, *d_lrec = nullptr;
    if ((rc = tmp.alloc(&d_face, (size_t)c->ns * n * (nv + 1))) || (rc = tmp.alloc(&d_lrec, (size_t)nv * n))) return rc;
    t.face = d_face;
    t.lrec = d_lrec;
    if (c->ph.viscous) {
        double *d_sl = nullptr, *d_gx = nullptr, *d_fgr = nullptr;
        if ((rc = tmp.alloc(&d_sl, (size_t)c->ns * n * nv)) ||
            (rc = tmp.alloc(&d_gx, (size_t)2 * dim * std::max(1, c->nb))) ||
            (rc = tmp.alloc(&d_fgr, (size_t)c->ns * n * (dim * dim + dim))))
            return rc;
        t.fgrad = d_fgr;
        t.slot = d_sl;
        t.gbx = d_gx;
        // the per-face geometry depends on the mesh and the scheme's slot geometry only: built on the
        // first adjoint call after ty_set_scheme and kept (the context's buffer, freed with the scheme)
        const bool fresh = c->d_adj_fgeo == nullptr;
        if (fresh && (rc = dalloc(c, &c->d_adj_fgeo, (size_t)4 * c->nf))) return rc;
        t.fgeo = c->d_adj_fgeo;
        if (fresh) HIPCHK(c, dim == 2 ? launch_adj_face_geom<2>(t, c->stream) : launch_adj_face_geom<3>(t, c->stream));
    }
    return TY_OK;
}

}  // namespace

int ty_residual_tangent(ty_ctx* c, double dt, const double* du, const double* ds, const double* dvol,
                        const double* dq, double* dres)
{
    if (!c || !dres) return TY_ERR_ARG;
    AdArgs t;
    int rc;
    DevTmp tmp(c);
    if (c->ph.viscous && ds)
        return fail(c, TY_ERR_UNSUPPORTED, "tangent of the viscous residual: area-vector directions not provided");
    if (c->mixed)
        return fail(c, TY_ERR_UNSUPPORTED, "tangent residual: FP64 schemes only (it differentiates the FP64 residual, "
                                           "which a mixed-precision context does not evaluate)");
    if ((rc = ad_prepare(c, dt, t, tmp))) return rc;
    const int n = c->n, nf = c->nf, nv = c->nv;
    double *d_du = nullptr, *d_ds = nullptr, *d_dv = nullptr, *d_dq = nullptr, *d_out = nullptr;
    std::vector<double> h;
    if (du) {
        host_to_dev_vec(c, du, h);
        if ((rc = tmp.alloc(&d_du, h.size()))) return rc;
        HIPCHK(c, hipMemcpy(d_du, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (ds) {
        h.assign((size_t)3 * nf, 0.0);
        for (int f = 0; f < nf; ++f)
            for (int e = 0; e < 3; ++e) h[(size_t)e * nf + f] = ds[(size_t)3 * c->fsrc[f] + e];
        if ((rc = tmp.alloc(&d_ds, h.size()))) return rc;
        HIPCHK(c, hipMemcpy(d_ds, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (dvol) {
        h.assign(n, 0.0);
        for (int i = 0; i < n; ++i) h[i] = dvol[c->perm[i]];
        if ((rc = tmp.alloc(&d_dv, h.size()))) return rc;
        HIPCHK(c, hipMemcpy(d_dv, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (dq) {
        const int nbc = (int)c->bocos.size();
        h.assign((size_t)std::max(1, nbc) * NVMAX, 0.0);
        for (int b = 0; b < nbc; ++b)
            for (int v = 0; v < NVMAX; ++v) h[(size_t)b * NVMAX + v] = dq[(size_t)b * 5 + v];
        if ((rc = tmp.alloc(&d_dq, h.size()))) return rc;
        HIPCHK(c, hipMemcpy(d_dq, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if ((rc = tmp.alloc(&d_out, (size_t)nv * n))) return rc;
    t.du = d_du;
    t.ds = d_ds;
    t.dvol = d_dv;
    t.dq = d_dq;
    t.out = d_out;
    if (c->ph.viscous) {
        double* d_dg = nullptr;
        if ((rc = tmp.alloc(&d_dg, (size_t)nv * c->dim * n))) return rc;
        t.dgrad = d_dg;
    }
    {
        EventScope ev(c, c->ev_tan);
        HIPCHK(c, c->dim == 2 ? launch_res_tangent<2>(t, c->stream) : launch_res_tangent<3>(t, c->stream));
    }
    h.assign((size_t)nv * n, 0.0);
    HIPCHK(c, hipMemcpyAsync(h.data(), d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_to_host_vec(c, h, dres);
    return TY_OK;
}

int ty_residual_adjoint(ty_ctx* c, double dt, const double* lam, double* gu, double* gs, double* gvol, double* gq)
{
    if (!c || !lam || !gu) return TY_ERR_ARG;
    AdArgs t;
    int rc;
    DevTmp tmp(c);
    if (c->ph.viscous && gs)
        return fail(c, TY_ERR_UNSUPPORTED, "adjoint of the viscous residual: area-vector sensitivities not provided");
    if (c->ph.viscous && c->ph.sch_hyp != 25 && c->ph.sch_hyp != 26)
        return fail(c, TY_ERR_UNSUPPORTED, "adjoint of the viscous residual: HLLC/HLLCB fluxes only");
    if (c->mixed)
        return fail(c, TY_ERR_UNSUPPORTED, "adjoint residual: FP64 schemes only (it transposes the FP64 residual's "
                                           "derivative, which a mixed-precision context does not evaluate)");
    if ((rc = ad_prepare(c, dt, t, tmp))) return rc;
    const int n = c->n, nf = c->nf, nb = c->nb, nv = c->nv;
    double *d_lam = nullptr, *d_gu = nullptr, *d_gs = nullptr, *d_gv = nullptr, *d_gq = nullptr;
    std::vector<double> h;
    host_to_dev_vec(c, lam, h);
    if ((rc = tmp.alloc(&d_lam, h.size())) || (rc = tmp.alloc(&d_gu, h.size()))) return rc;
    HIPCHK(c, hipMemcpy(d_lam, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = alloc_adjoint_work(c, t, tmp))) return rc;
    {
        // lambda dt / V as one record per cell (the HLLC face kernels), marshalled with the upload
        std::vector<double> lr((size_t)nv * n);
        parallel_for(n, [&](int i) {
            const double k = dt / c->vol[i];
            for (int v = 0; v < nv; ++v) lr[(size_t)i * nv + v] = h[(size_t)v * n + i] * k;
        });
        HIPCHK(c, hipMemcpy(t.lrec, lr.data(), lr.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (gs && (rc = tmp.alloc(&d_gs, (size_t)3 * nf))) return rc;
    if (gvol && (rc = tmp.alloc(&d_gv, (size_t)n))) return rc;
    if (gq && (rc = tmp.alloc(&d_gq, (size_t)nv * std::max(1, nb)))) return rc;
    t.lam = d_lam;
    t.gu = d_gu;
    t.gs = d_gs;
    t.gvol = d_gv;
    t.gq = d_gq;
    {
        EventScope ev(c, c->ev_adj);
        HIPCHK(c, c->dim == 2 ? launch_res_adjoint<2>(t, c->stream) : launch_res_adjoint<3>(t, c->stream));
    }
    h.assign((size_t)nv * n, 0.0);
    HIPCHK(c, hipMemcpyAsync(h.data(), d_gu, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> hs(gs ? (size_t)3 * nf : 0), hv(gvol ? (size_t)n : 0), hq(gq ? (size_t)nv * nb : 0);
    if (gs) HIPCHK(c, hipMemcpyAsync(hs.data(), d_gs, hs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (gvol) HIPCHK(c, hipMemcpyAsync(hv.data(), d_gv, hv.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (gq && nb) HIPCHK(c, hipMemcpyAsync(hq.data(), d_gq, hq.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_to_host_vec(c, h, gu);
    if (gs) {
        // caller faces [nface][3]; faces whose geometry the device does not use (the far face of
        // a periodic pair) get zero
        std::fill(gs, gs + (size_t)3 * c->nface_caller, 0.0);
        for (int f = 0; f < nf; ++f)
            for (int e = 0; e < 3; ++e) gs[(size_t)3 * c->fsrc[f] + e] = hs[(size_t)e * nf + f];
    }
    if (gvol)
        for (int i = 0; i < n; ++i) gvol[c->perm[i]] = hv[i];
    if (gq) {
        const int nbc = (int)c->bocos.size();
        std::fill(gq, gq + (size_t)5 * nbc, 0.0);
        for (int k = 0; k < nb; ++k)  // fixed order: device boundary faces
            for (int v = 0; v < nv; ++v) gq[(size_t)5 * c->bboco[k] + v] += hq[(size_t)v * nb + k];
    }
    return TY_OK;
}

// ---------------------------------------------------------------------------
// Unsteady adjoint: reverse sweep over RK steps with device-resident checkpoints.

namespace {

// Binomial checkpointing (Griewank, Optim. Methods Softw. 1 (1992)): with `snaps` free
// snapshots over `steps` steps the repetition number r is the smallest with
// C(snaps + r, snaps) >= steps; the right part of the range keeps C(snaps - 1 + r, snaps - 1) steps.
double binom(int n, int k)
{
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

int tv_split(int steps, int snaps)
{
    if (steps < 2) return 1;
    int r = 0;
    while (binom(snaps + r, snaps) < steps) ++r;
    const double right = snaps >= 1 ? binom(snaps - 1 + r, snaps - 1) : 1.0;
    const double left = steps - right;
    return (int)std::min<double>(std::max<double>(1.0, left), steps - 1);
}

struct Unsteady {
    ty_ctx* c = nullptr;
    long len = 0;                // nv n
    int ns = 0;                  // RK stages
    const double* dts = nullptr;  // [nsteps] device
    std::vector<double*> pool;   // snapshot buffers
    double* work = nullptr;      // recomputation state for ranges without snapshots
    double* stU = nullptr;       // [ns - 1][len] stage states 2..ns
    double* stR = nullptr;       // [ns][len] stage residuals
    double* nu = nullptr;        // [ns][len] stage-state adjoints
    double* mu = nullptr;        // [len] stage-residual adjoint
    double* rec = nullptr;       // [len] stage state as residual-kernel records
    double* gq_tmp = nullptr;    // [nv][nb] inlet sensitivities of one residual adjoint
    double* gq_acc = nullptr;    // [nv][nb] summed
    void* us64[2] = {nullptr, nullptr};  // FP64 stage-state records of a mixed-precision context
    AdArgs t;
    int live = 0, fwd = 0, adj = 0, maxlive = 0;

    double coef(int i, int j) const { return c->rk[(size_t)(i - 1) * ns + (j - 1)]; }

    int set_dt(int k)
    {
        HIPCHK(c, hipMemcpyAsync(c->d_dt, dts + k, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        return TY_OK;
    }

    // U <- state after step k by the FP64 RK path (the map the reverse sweep linearises; a
    // mixed-precision context runs it on FP64 stage records of its own)
    int advance(double* U, int k)
    {
        int rc;
        if ((rc = set_dt(k))) return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_u0, U, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        void* keep[2] = {c->d_us[0], c->d_us[1]};
        if (c->mixed) {
            c->d_us[0] = us64[0];
            c->d_us[1] = us64[1];
        }
        c->cur = 0;
        hipError_t e = launch_to_stage<double>(c->d_u0, (double*)c->d_us[0], c->nv, c->n, c->gas.gamma, c->stream);
        if (e == hipSuccess) e = c->dim == 2 ? rk_stages<2, double>(c, c->geo_d) : rk_stages<3, double>(c, c->geo_d);
        c->d_us[0] = keep[0];
        c->d_us[1] = keep[1];
        HIPCHK(c, e);
        HIPCHK(c, hipMemcpyAsync(U, c->d_u0, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        ++fwd;
        return TY_OK;
    }

    // out = (dR/dU at U)^T m, FP64; inlet sensitivities added to gq_acc.  The HLLC kernels read m as
    // the lambda dt / V records in t.lrec, which the caller has filled from the same combination.
    int adjoint_apply(const double* U, const double* m, double* out)
    {
        AdArgs a = t;
        a.k.us = U;
        HIPCHK(c, launch_to_stage<double>(U, rec, c->nv, c->n, c->gas.gamma, c->stream));
        a.qrec = rec;
        if (c->ph.viscous) {
            KArgs<double> g = make_args<double>(c, c->geo_d);
            g.us = rec;
            g.qf_out = nullptr;
            g.vrec = 0;  // full gradients for the adjoint kernels
            g.vrec_out = const_cast<double*>(t.vgrad);  // and their viscous records
            const hipError_t e = c->dim == 2 ? launch_gradient<2, double>(g, c->stream)
                                             : launch_gradient<3, double>(g, c->stream);
            HIPCHK(c, e);
            a.k.grad = (const double*)c->work_grad[0];
        }
        a.lam = m;
        a.gu = out;
        a.gq = gq_tmp;
        HIPCHK(c, c->dim == 2 ? launch_res_adjoint<2>(a, c->stream) : launch_res_adjoint<3>(a, c->stream));
        if (gq_acc) HIPCHK(c, accumulate<double>(gq_acc, gq_tmp, (long)c->nv * c->nb, c->stream));
        return TY_OK;
    }

    // lam <- (dU_{k+1}/dU_k)^T lam, step k starting from Uk (treelevel_rungekutta.f90:60-105:
    // U_s = U_k + sum_{j<s} coef(s-1, j) R_j, U_{k+1} = U_k + sum_j coef(ns, j) R_j)
    int step_adjoint(int k, const double* Uk, double* lam)
    {
        int rc;
        if ((rc = set_dt(k))) return rc;
        const double* Us[MAX_STAGES];
        Us[0] = Uk;
        for (int s = 1; s <= ns; ++s) {
            HIPCHK(c, launch_to_stage<double>(Us[s - 1], rec, c->nv, c->n, c->gas.gamma, c->stream));
            HIPCHK(c, eval_residual<double>(c, rec, stR + (size_t)(s - 1) * len, nullptr, nullptr));
            if (s == ns) break;
            Comb cb;
            memset(&cb, 0, sizeof(cb));
            cb.v[0] = Uk;
            cb.c[0] = 1.0;
            cb.m = 1;
            for (int j = 1; j <= s; ++j) {
                cb.v[cb.m] = stR + (size_t)(j - 1) * len;
                cb.c[cb.m] = coef(s, j);
                ++cb.m;
            }
            double* out = stU + (size_t)(s - 1) * len;
            HIPCHK(c, launch_comb(out, cb, len, c->stream));
            Us[s] = out;
        }
        // reverse: mu_s = coef(ns, s) lam + sum_{m>s} coef(m-1, s) nu_m, nu_s = (dR/dU at U_s)^T mu_s
        for (int s = ns; s >= 1; --s) {
            Comb cb;
            memset(&cb, 0, sizeof(cb));
            cb.v[0] = lam;
            cb.c[0] = coef(ns, s);
            cb.m = 1;
            for (int m = s + 1; m <= ns; ++m) {
                cb.v[cb.m] = nu + (size_t)(m - 1) * len;
                cb.c[cb.m] = coef(m - 1, s);
                ++cb.m;
            }
            // the HLLC kernels take mu as lambda dt / V records, the dual-number kernel as a vector
            if (c->ph.sch_hyp == 25 || c->ph.sch_hyp == 26)
                HIPCHK(c, launch_comb_lrec(t.lrec, cb, c->d_vol, c->d_dt, c->nv, c->n, c->stream));
            else
                HIPCHK(c, launch_comb(mu, cb, len, c->stream));
            if ((rc = adjoint_apply(Us[s - 1], mu, nu + (size_t)(s - 1) * len))) return rc;
        }
        // lam_k = lam + sum_s nu_s
        Comb cb;
        memset(&cb, 0, sizeof(cb));
        cb.v[0] = lam;
        cb.c[0] = 1.0;
        cb.m = 1;
        for (int s = 1; s <= ns; ++s) {
            cb.v[cb.m] = nu + (size_t)(s - 1) * len;
            cb.c[cb.m] = 1.0;
            ++cb.m;
        }
        HIPCHK(c, launch_comb(lam, cb, len, c->stream));
        ++adj;
        return TY_OK;
    }

    // lam: adjoint at step b on entry, at step a on return; Ua: state at step a (kept)
    int treeverse(int a, int b, const double* Ua, double* lam, int snaps)
    {
        int rc;
        if (b - a == 1) return step_adjoint(a, Ua, lam);
        if (snaps == 0) {
            for (int k = b - 1; k >= a; --k) {
                HIPCHK(c, hipMemcpyAsync(work, Ua, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
                for (int j = a; j < k; ++j)
                    if ((rc = advance(work, j))) return rc;
                if ((rc = step_adjoint(k, work, lam))) return rc;
            }
            return TY_OK;
        }
        const int m = a + tv_split(b - a, snaps);
        // live <= min(snaps of the call, nsteps - 1) = pool.size(): each level holds one more
        // snapshot with one snapshot less and at least one step less to cover
        if (live >= (int)pool.size()) return fail(c, TY_ERR_ARG, "unsteady adjoint: snapshot pool exhausted");
        double* Um = pool[live++];
        maxlive = std::max(maxlive, live);
        HIPCHK(c, hipMemcpyAsync(Um, Ua, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        for (int j = a; j < m; ++j)
            if ((rc = advance(Um, j))) return rc;
        if ((rc = treeverse(m, b, Um, lam, snaps - 1))) return rc;
        --live;
        return treeverse(a, m, Ua, lam, snaps);
    }
};

}  // namespace

int ty_unsteady_adjoint(ty_ctx* c, int32_t nsteps, const double* dts, int32_t snaps, const double* lamN,
                        double* lam0, double* gq, ty_unsteady_stats* stats)
{
    if (!c || nsteps < 0 || snaps < 0 || (nsteps > 0 && !dts) || !lamN || !lam0) return TY_ERR_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    if (c->ph.viscous && c->ph.sch_hyp != 25 && c->ph.sch_hyp != 26)
        return fail(c, TY_ERR_UNSUPPORTED, "adjoint of the viscous residual: HLLC/HLLCB fluxes only");
    if (c->ext_bc)
        return fail(c, TY_ERR_UNSUPPORTED, "unsteady adjoint: wall and subsonic-inlet boundaries not provided");
    DevTmp tmp(c);
    Unsteady u;
    int rc;
    if ((rc = ad_prepare(c, nsteps > 0 ? dts[0] : 0.0, u.t, tmp))) return rc;
    if ((rc = ensure_work(c, 0))) return rc;
    u.c = c;
    u.ns = c->sch.nstage;
    const int n = c->n, nv = c->nv, nb = c->nb;
    u.len = (long)nv * n;
    const size_t len = (size_t)u.len;
    double *d_dts = nullptr, *U0 = nullptr, *lam = nullptr;
    if ((rc = tmp.alloc(&d_dts, (size_t)std::max(1, nsteps))) || (rc = tmp.alloc(&U0, len)) ||
        (rc = tmp.alloc(&lam, len)) || (rc = tmp.alloc(&u.work, len)) ||
        (rc = tmp.alloc(&u.stU, (size_t)std::max(1, u.ns - 1) * len)) || (rc = tmp.alloc(&u.stR, (size_t)u.ns * len)) ||
        (rc = tmp.alloc(&u.nu, (size_t)u.ns * len)) || (rc = tmp.alloc(&u.mu, len)) || (rc = tmp.alloc(&u.rec, len)) ||
        (rc = tmp.alloc(&u.gq_tmp, (size_t)nv * std::max(1, nb))))
        return rc;
    // treeverse holds at most min(snaps, nsteps - 1) snapshots at once
    const int nsnap = std::min(snaps, std::max(0, nsteps - 1));
    u.pool.assign(nsnap, nullptr);
    for (int k = 0; k < nsnap; ++k)
        if ((rc = tmp.alloc(&u.pool[k], len))) return rc;
    if (c->mixed)
        for (int k = 0; k < 2; ++k)
            if ((rc = tmp.alloc((double**)&u.us64[k], len))) return rc;
    if (gq) {
        if ((rc = tmp.alloc(&u.gq_acc, (size_t)nv * std::max(1, nb)))) return rc;
        HIPCHK(c, hipMemsetAsync(u.gq_acc, 0, (size_t)nv * std::max(1, nb) * sizeof(double), c->stream));
    }
    if ((rc = alloc_adjoint_work(c, u.t, tmp))) return rc;
    u.t.gs = nullptr;
    u.t.gvol = nullptr;
    if (nsteps > 0)
        HIPCHK(c, hipMemcpy(d_dts, dts, (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice));
    u.dts = d_dts;
    std::vector<double> h;
    host_to_dev_vec(c, lamN, h);
    HIPCHK(c, hipMemcpy(lam, h.data(), len * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpyAsync(U0, c->d_u0, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (nsteps > 0 && (rc = u.treeverse(0, nsteps, U0, lam, snaps))) return rc;
    // the state is U_0 again
    HIPCHK(c, hipMemcpyAsync(c->d_u0, U0, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->cur = 0;
    hipError_t e = c->mixed ? launch_to_stage<float>(c->d_u0, (float*)c->d_us[0], nv, n, c->gas.gamma, c->stream)
                            : launch_to_stage<double>(c->d_u0, (double*)c->d_us[0], nv, n, c->gas.gamma, c->stream);
    HIPCHK(c, e);
    HIPCHK(c, hipMemcpyAsync(h.data(), lam, len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> hq(gq ? (size_t)nv * nb : 0);
    if (gq && nb) HIPCHK(c, hipMemcpyAsync(hq.data(), u.gq_acc, hq.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_to_host_vec(c, h, lam0);
    if (gq) {
        const int nbc = (int)c->bocos.size();
        std::fill(gq, gq + (size_t)5 * nbc, 0.0);
        for (int k = 0; k < nb; ++k)
            for (int v = 0; v < nv; ++v) gq[(size_t)5 * c->bboco[k] + v] += hq[(size_t)v * nb + k];
    }
    if (stats) {
        stats->forward_steps = u.fwd;
        stats->adjoint_steps = u.adj;
        stats->max_snapshots = u.maxlive;
        stats->snapshots_allocated = nsnap;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return TY_OK;
}

// ---------------------------------------------------------------------------
// the reference's assembled implicit step: calc_rhs with face Jacobians, build_implicit_bdlu,
// bdlu_gmres, update_field (tstep_implicit.f90, implicit_solve.f90:39-84)

int ty_implicit_bdlu_step(ty_ctx* c, double dt, int32_t nkrylov, double maxres, int32_t max_it, int32_t* info)
{
    if (!c || nkrylov < 1 || max_it < 0 || !info) return TY_ERR_ARG;
    if (!c->state_ok) return fail(c, TY_ERR_ARG, "no state");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_work(c, 0))) return rc;
    DevTmp tmp(c);
    const int n = c->n, nv = c->nv, ns = c->ns;
    const long ng = (long)c->nb + 2L * c->nper;
    const long N = n + ng, L = (long)nv * N;
    const int m = nkrylov;
    double *stage, *res, *D, *B, *p1, *sol, *r1, *w1, *V, *part, *hd;
    int* tgt;
    if ((rc = tmp.alloc(&stage, (size_t)nv * n)) || (rc = tmp.alloc(&res, (size_t)nv * n)) ||
        (rc = tmp.alloc(&D, (size_t)nv * nv * n)) || (rc = tmp.alloc(&B, (size_t)ns * nv * nv * n)) ||
        (rc = tmp.alloc(&tgt, (size_t)ns * n)) || (rc = tmp.alloc(&p1, (size_t)L)) || (rc = tmp.alloc(&sol, (size_t)L)) ||
        (rc = tmp.alloc(&r1, (size_t)L)) || (rc = tmp.alloc(&w1, (size_t)L)) || (rc = tmp.alloc(&V, (size_t)m * L)) ||
        (rc = tmp.alloc(&part, (size_t)1024 * MAXK)) || (rc = tmp.alloc(&hd, (size_t)MAXK)))
        return rc;
    c->h_dt = dt;
    HIPCHK(c, hipMemcpyAsync(c->d_dt, &c->h_dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(tgt, 0xff, (size_t)ns * n * sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(p1, 0, (size_t)L * sizeof(double), c->stream));
    // residual of every unknown (interior rows -dt/V sum F.A; ghost rows their boundary-flux sums),
    // diagonal blocks with ns_bocojacobian, per-slot coupling blocks
    HIPCHK(c, (fill_state<double, double>(stage, c->d_u0, 0.0, c->d_u0, nv, n, c->gas.gamma, c->stream)));
    KArgs<double> a = make_args<double>(c, c->geo_d);
    a.us = stage;
    a.mode = 1;
    a.res_out = res;
    a.diag_out = D;
    a.dinv_out = nullptr;
    a.bdlu_B = B;
    a.bdlu_tgt = tgt;
    a.bdlu_grhs = p1 + n;
    a.bdlu_ld = N;
    a.nb_dev = c->nb;
    a.per_base = c->per_base;
    HIPCHK(c, (c->dim == 2 ? eval_kernels<2, double>(c, a) : eval_kernels<3, double>(c, a)));
    for (int v = 0; v < nv; ++v)
        HIPCHK(c, hipMemcpyAsync(p1 + (size_t)v * N, res + (size_t)v * n, (size_t)n * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
    auto Ax = [&](const double* x, double* y) { return bdlu_apply(nv, D, B, tgt, ns, n, N, x, y, c->stream); };
    auto dot = [&](const double* x, const double* y, double* out) -> int {
        HIPCHK(c, multidot<double>(x, y, 0, 1, L, part, hd, c->stream));
        HIPCHK(c, hipMemcpyAsync(out, hd, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TY_OK;
    };
    // bdlu_gmres.f90:55-212
    std::vector<double> hh((size_t)(m + 1) * m), ss(m + 1), qc(m), qs(m), yy(m);
    auto HH = [&](int i, int j) -> double& { return hh[(size_t)(j - 1) * (m + 1) + (i - 1)]; };
    HIPCHK(c, bdlu_x0(D, p1, sol, nv, n, N, c->stream));
    double ref = 0.0, pp = 0.0;
    HIPCHK(c, asum(sol, L, part, hd, c->stream));
    HIPCHK(c, hipMemcpyAsync(&ref, hd, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = dot(p1, p1, &pp))) return rc;
    const double normp1 = std::sqrt(pp);
    double errgmres = std::numeric_limits<double>::max();
    int nit = 0;
    while (errgmres >= ref * maxres && nit <= max_it * 2) {
        HIPCHK(c, Ax(sol, r1));
        HIPCHK(c, pmr(r1, p1, L, c->stream));
        double rr;
        if ((rc = dot(r1, r1, &rr))) return rc;
        const double beta = std::sqrt(rr);
        HIPCHK(c, (scale<double, double>(V, r1, 1.0 / beta, L, c->stream)));
        std::fill(ss.begin(), ss.end(), 0.0);
        ss[0] = beta;
        std::fill(hh.begin(), hh.end(), 0.0);
        int jk;
        for (jk = 1; jk <= m; ++jk) {
            HIPCHK(c, Ax(V + (size_t)(jk - 1) * L, w1));
            for (int ik = 1; ik <= jk; ++ik) {
                const double* vi = V + (size_t)(ik - 1) * L;
                if ((rc = dot(w1, vi, &HH(ik, jk)))) return rc;
                HIPCHK(c, axpy(w1, vi, -HH(ik, jk), L, c->stream));
            }
            double ww;
            if ((rc = dot(w1, w1, &ww))) return rc;
            HH(jk + 1, jk) = std::sqrt(ww);
            for (int ik = 1; ik <= jk - 1; ++ik) {
                const double t = HH(ik, jk);
                HH(ik, jk) = qc[ik - 1] * t + qs[ik - 1] * HH(ik + 1, jk);
                HH(ik + 1, jk) = -qs[ik - 1] * t + qc[ik - 1] * HH(ik + 1, jk);
            }
            if (HH(jk + 1, jk) > std::numeric_limits<double>::epsilon() * 10 * normp1) {
                if (jk < m) HIPCHK(c, (scale<double, double>(V + (size_t)jk * L, w1, 1.0 / HH(jk + 1, jk), L, c->stream)));
                if (HH(jk, jk) == 0) break;
                const double t = std::sqrt(HH(jk, jk) * HH(jk, jk) + HH(jk + 1, jk) * HH(jk + 1, jk));
                qc[jk - 1] = HH(jk, jk) / t;
                qs[jk - 1] = HH(jk + 1, jk) / t;
                HH(jk, jk) = t;
                HH(jk + 1, jk) = 0.0;
                ss[jk] = -qs[jk - 1] * ss[jk - 1];
                ss[jk - 1] = qc[jk - 1] * ss[jk - 1];
                if (std::fabs(ss[jk]) < maxres * std::fabs(ss[0])) break;
            } else {
                break;
            }
        }
        const int nk = std::min(jk, m);
        for (int j = nk; j >= 1; --j) {
            double t = ss[j - 1];
            for (int ik = j + 1; ik <= nk; ++ik) t -= HH(j, ik) * yy[ik - 1];
            yy[j - 1] = t / HH(j, j);
            HIPCHK(c, axpy(sol, V + (size_t)(j - 1) * L, yy[j - 1], L, c->stream));
        }
        HIPCHK(c, Ax(sol, r1));
        HIPCHK(c, pmr(r1, p1, L, c->stream));
        if ((rc = dot(r1, r1, &rr))) return rc;
        ss[nk] = std::sqrt(rr);
        errgmres = ss[nk] / normp1;
        ++nit;
    }
    *info = nit <= max_it ? nit - 1 : -1;
    // update_field.f90:38 on the interior unknowns, then the stage-state copy
    for (int v = 0; v < nv; ++v)
        HIPCHK(c, accumulate<double>(c->d_u0 + (size_t)v * n, sol + (size_t)v * N, n, c->stream));
    hipError_t e = c->mixed ? launch_to_stage<float>(c->d_u0, (float*)c->d_us[c->cur], nv, n, c->gas.gamma, c->stream)
                            : launch_to_stage<double>(c->d_u0, (double*)c->d_us[c->cur], nv, n, c->gas.gamma, c->stream);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TY_OK;
}

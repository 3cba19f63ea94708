// Reverse Cuthill-McKee order of the interior-cell graph on the device.
//
// The serial algorithm: a pseudo-peripheral start (a breadth-first search from the first unordered
// cell, then one from the lowest-degree cell of its last level, whose last level gives the start);
// from the start, a breadth-first queue in which each dequeued cell appends its not yet queued
// neighbours sorted by degree (stably, in adjacency order); one such pass per connected component
// (started from the lowest-numbered unordered cell); the concatenated queue reversed.
//
// Here each pass runs level by level.  A cell of the next level belongs to the first cell of the
// current level (in queue order) that reaches it, as in the serial queue, so its place is fixed by
// the key (parent's queue position, degree, index in the parent's adjacency list) -- without the
// degree for the plain searches.  Every neighbour of the level takes the minimum key of the
// parents that reach it (a 64-bit atomic minimum; the first taker also lists the cell), and one radix
// sort of the listed cells by key gives the level in queue order: the serial order, exactly.
// Adjacency lists are in edge order (a sort of (cell, edge) keys), as the serial version builds them.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "ty_internal.h"

namespace ty {

namespace {

constexpr unsigned long long NONE = ~0ull;
constexpr int KBITS = 12;  // degree and adjacency index fields of a key (degrees below 4096)

__global__ void k_edge_entries(int ne, const int* __restrict__ ea, const int* __restrict__ eb,
                               unsigned long long* __restrict__ key, int* __restrict__ dst, int* __restrict__ deg)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int a = ea[e], b = eb[e];
    key[2 * (size_t)e] = ((unsigned long long)(unsigned)a << 32) | (unsigned)e;
    dst[2 * (size_t)e] = b;
    key[2 * (size_t)e + 1] = ((unsigned long long)(unsigned)b << 32) | (unsigned)e;
    dst[2 * (size_t)e + 1] = a;
    atomicAdd(deg + a, 1);
    atomicAdd(deg + b, 1);
}

__global__ void k_fill(int n, unsigned long long* __restrict__ claim, int* __restrict__ mark, char* __restrict__ done)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    claim[i] = NONE;
    mark[i] = 0;
    done[i] = 0;
}

// one level: every unreached neighbour v of the cells q[lo..hi) takes the minimum key
__global__ void k_expand(int lo, int hi, int stamp, int cm, const int* __restrict__ q, const int* __restrict__ xadj,
                         const int* __restrict__ adj, const int* __restrict__ deg, const int* __restrict__ mark,
                         const char* __restrict__ done, unsigned long long* __restrict__ claim, int* __restrict__ cand,
                         int* __restrict__ ncand)
{
    const int p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi) return;
    const int u = q[p];
    for (int k = xadj[u]; k < xadj[u + 1]; ++k) {
        const int v = adj[k];
        if (done[v] || mark[v] == stamp) continue;
        const unsigned long long kidx = (unsigned long long)(k - xadj[u]);
        const unsigned long long key = cm ? ((unsigned long long)p << (2 * KBITS)) |
                                                ((unsigned long long)deg[v] << KBITS) | kidx
                                          : ((unsigned long long)p << KBITS) | kidx;
        const unsigned long long old = atomicMin(claim + v, key);
        if (old == NONE) cand[atomicAdd(ncand, 1)] = v;
    }
}

__global__ void k_claim_keys(int m, const int* __restrict__ cand, const unsigned long long* __restrict__ claim,
                             unsigned long long* __restrict__ key)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) key[i] = claim[cand[i]];
}

// the sorted level appended to the queue at t
__global__ void k_place(int m, int t, int stamp, const int* __restrict__ sorted, int* __restrict__ q,
                        int* __restrict__ mark, unsigned long long* __restrict__ claim)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int v = sorted[i];
    q[t + i] = v;
    mark[v] = stamp;
    claim[v] = NONE;
}

__global__ void k_start(int s, int t, int stamp, int* __restrict__ q, int* __restrict__ mark)
{
    q[t] = s;
    mark[s] = stamp;
}

// lowest (degree, queue position) over q[lo..hi)
__global__ void k_min_degree(int lo, int hi, const int* __restrict__ q, const int* __restrict__ deg,
                             unsigned long long* __restrict__ best)
{
    const int p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi) return;
    atomicMin(best, ((unsigned long long)deg[q[p]] << 40) | (unsigned long long)p);
}

__global__ void k_mark_done(int lo, int hi, const int* __restrict__ q, char* __restrict__ done)
{
    const int p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p < hi) done[q[p]] = 1;
}

__global__ void k_first_undone(int n, const char* __restrict__ done, int* __restrict__ first)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !done[i]) atomicMin(first, i);
}

inline unsigned nblk(long n) { return (unsigned)std::max<long>(1, (n + 255) / 256); }

struct Bufs {
    std::vector<void*> p;
    ~Bufs()
    {
        for (void* x : p) (void)hipFree(x);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t count)
    {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};

#define OCHK(expr)                              \
    do {                                        \
        const hipError_t e_ = (expr);           \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

}  // namespace

hipError_t rcm_order_device(int n, const std::vector<int>& ea, const std::vector<int>& eb, std::vector<int>& order,
                            bool& degree_ok, hipStream_t st)
{
    degree_ok = true;
    order.assign(n, 0);
    if (n <= 0) return hipSuccess;
    const int ne = (int)ea.size();
    const int nd = 2 * ne;
    Bufs b;
    int *d_ea, *d_eb, *d_dst, *d_adj, *d_deg, *d_xadj, *d_mark, *d_q, *d_cand, *d_sorted, *d_cnt, *d_first;
    unsigned long long *d_key, *d_skey, *d_claim, *d_ckey, *d_cskey, *d_best;
    char* d_done;
    OCHK(b.alloc(&d_ea, ne));
    OCHK(b.alloc(&d_eb, ne));
    OCHK(b.alloc(&d_dst, nd));
    OCHK(b.alloc(&d_adj, nd));
    OCHK(b.alloc(&d_key, nd));
    OCHK(b.alloc(&d_skey, nd));
    OCHK(b.alloc(&d_deg, n));
    OCHK(b.alloc(&d_xadj, n + 1));
    OCHK(b.alloc(&d_mark, n));
    OCHK(b.alloc(&d_q, n));
    OCHK(b.alloc(&d_cand, n));
    OCHK(b.alloc(&d_sorted, n));
    OCHK(b.alloc(&d_claim, n));
    OCHK(b.alloc(&d_ckey, n));
    OCHK(b.alloc(&d_cskey, n));
    OCHK(b.alloc(&d_done, n));
    OCHK(b.alloc(&d_cnt, 1));
    OCHK(b.alloc(&d_first, 1));
    OCHK(b.alloc(&d_best, 1));
    if (ne) {
        OCHK(hipMemcpyAsync(d_ea, ea.data(), sizeof(int) * ne, hipMemcpyHostToDevice, st));
        OCHK(hipMemcpyAsync(d_eb, eb.data(), sizeof(int) * ne, hipMemcpyHostToDevice, st));
    }
    OCHK(hipMemsetAsync(d_deg, 0, sizeof(int) * n, st));
    k_fill<<<nblk(n), 256, 0, st>>>(n, d_claim, d_mark, d_done);
    if (ne) k_edge_entries<<<nblk(ne), 256, 0, st>>>(ne, d_ea, d_eb, d_key, d_dst, d_deg);
    OCHK(hipGetLastError());
    // adjacency lists in edge order; xadj = exclusive scan of the degrees
    size_t t1 = 0, t2 = 0, t3 = 0;
    OCHK(rocprim::radix_sort_pairs(nullptr, t1, d_key, d_skey, d_dst, d_adj, nd, 0, 64, st));
    OCHK(rocprim::exclusive_scan(nullptr, t2, d_deg, d_xadj, 0, (size_t)n, rocprim::plus<int>(), st));
    OCHK(rocprim::radix_sort_pairs(nullptr, t3, d_ckey, d_cskey, d_cand, d_sorted, n, 0, 64, st));
    void* tmp = nullptr;
    const size_t tmax = std::max(t1, std::max(t2, t3));
    OCHK(b.alloc((char**)&tmp, tmax));
    if (nd) OCHK(rocprim::radix_sort_pairs(tmp, t1, d_key, d_skey, d_dst, d_adj, nd, 0, 64, st));
    OCHK(rocprim::exclusive_scan(tmp, t2, d_deg, d_xadj, 0, (size_t)n, rocprim::plus<int>(), st));
    std::vector<int> hdeg(n);
    OCHK(hipMemcpyAsync(hdeg.data(), d_deg, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    OCHK(hipStreamSynchronize(st));
    int maxdeg = 0;
    for (int i = 0; i < n; ++i) maxdeg = std::max(maxdeg, hdeg[i]);
    if (maxdeg >= (1 << KBITS)) {
        degree_ok = false;
        return hipSuccess;
    }
    OCHK(hipMemcpy(d_xadj + n, &nd, sizeof(int), hipMemcpyHostToDevice));  // xadj[n] = all entries

    int stamp = 0;
    // one search from s with the queue at q[t0..): returns the end of the queue; its levels'
    // bounds in lv (lv.back() = end)
    std::vector<int> lv;
    auto search = [&](int s, int t0, int cm, int& tend) -> hipError_t {
        ++stamp;
        k_start<<<1, 1, 0, st>>>(s, t0, stamp, d_q, d_mark);
        OCHK(hipGetLastError());
        lv.assign(1, t0);
        int lo = t0, hi = t0 + 1;
        while (true) {
            lv.push_back(hi);
            OCHK(hipMemsetAsync(d_cnt, 0, sizeof(int), st));
            k_expand<<<nblk(hi - lo), 256, 0, st>>>(lo, hi, stamp, cm, d_q, d_xadj, d_adj, d_deg, d_mark, d_done,
                                                      d_claim, d_cand, d_cnt);
            OCHK(hipGetLastError());
            int m = 0;
            OCHK(hipMemcpyAsync(&m, d_cnt, sizeof(int), hipMemcpyDeviceToHost, st));
            OCHK(hipStreamSynchronize(st));
            if (m == 0) break;
            k_claim_keys<<<nblk(m), 256, 0, st>>>(m, d_cand, d_claim, d_ckey);
            size_t tb = 0;
            OCHK(rocprim::radix_sort_pairs(nullptr, tb, d_ckey, d_cskey, d_cand, d_sorted, m, 0, 64, st));
            if (tb > tmax) return hipErrorInvalidValue;
            OCHK(rocprim::radix_sort_pairs(tmp, tb, d_ckey, d_cskey, d_cand, d_sorted, m, 0, 64, st));
            k_place<<<nblk(m), 256, 0, st>>>(m, hi, stamp, d_sorted, d_q, d_mark, d_claim);
            OCHK(hipGetLastError());
            lo = hi;
            hi += m;
        }
        tend = hi;
        return hipSuccess;
    };
    // the lowest-degree cell of the last level (queue order breaks ties), or s if none is lower
    auto far_end = [&](int s, int t0, int& out) -> hipError_t {
        int tend = 0;
        OCHK(search(s, t0, 0, tend));
        const int llo = lv[lv.size() - 2], lhi = lv.back();
        const unsigned long long init = NONE;
        OCHK(hipMemcpyAsync(d_best, &init, sizeof(init), hipMemcpyHostToDevice, st));
        k_min_degree<<<nblk(lhi - llo), 256, 0, st>>>(llo, lhi, d_q, d_deg, d_best);
        OCHK(hipGetLastError());
        unsigned long long best = 0;
        OCHK(hipMemcpyAsync(&best, d_best, sizeof(best), hipMemcpyDeviceToHost, st));
        OCHK(hipStreamSynchronize(st));
        const int bdeg = (int)(best >> 40);
        int node = s;
        if (bdeg < hdeg[s]) OCHK(hipMemcpy(&node, d_q + (int)(best & ((1ull << 40) - 1)), sizeof(int), hipMemcpyDeviceToHost));
        out = node;
        return hipSuccess;
    };
    int placed = 0, scan = 0;
    while (placed < n) {
        // the lowest-numbered cell not yet ordered
        const int big = n;
        OCHK(hipMemcpyAsync(d_first, &big, sizeof(int), hipMemcpyHostToDevice, st));
        k_first_undone<<<nblk(n), 256, 0, st>>>(n, d_done, d_first);
        OCHK(hipGetLastError());
        OCHK(hipMemcpyAsync(&scan, d_first, sizeof(int), hipMemcpyDeviceToHost, st));
        OCHK(hipStreamSynchronize(st));
        int s1 = 0, s2 = 0, tend = 0;
        OCHK(far_end(scan, placed, s1));
        OCHK(far_end(s1, placed, s2));
        OCHK(search(s2, placed, 1, tend));
        k_mark_done<<<nblk(tend - placed), 256, 0, st>>>(placed, tend, d_q, d_done);
        OCHK(hipGetLastError());
        placed = tend;
    }
    OCHK(hipMemcpyAsync(order.data(), d_q, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    OCHK(hipStreamSynchronize(st));
    std::reverse(order.begin(), order.end());
    return hipSuccess;
}

}  // namespace ty

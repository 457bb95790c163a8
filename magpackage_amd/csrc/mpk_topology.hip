// Mesh topology, DOF map, sparse() and the CSC plan on gfx950.
//
// Everything here is integer/byte work built on one primitive: a stable LSD
// radix sort of 64-bit keys (rocPRIM), followed by head flags, scans and
// scatters.  Stability is what makes the results match the reference's
// sorted-unique semantics and the triplet-order summation bit for bit.
//
//   mpk_generate_edges     src/Mesh.m:114-138 (unique(sort(...),'rows'), ic, signs)
//   mpk_distribute_dofs    src/DofHandler.m:72-105
//   mpk_sparse_coo_to_csc  src/Assembler.m:37-38 (I>0 & J>0 filter + sparse())
//   mpk_plan_create        the same sparse() pattern, built once, plus the
//                          per-column incidence lists used by the CSC kernels
#include <cstdlib>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "mpk_plan.hpp"

namespace mpk {
namespace {

constexpr int kB = 256;

unsigned bits_for(uint64_t maxval) {  // bits needed to represent values <= maxval
    unsigned b = 0;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b == 0 ? 1 : b;
}

// Stable sort of (uint64 key, uint32 value) pairs over bits [0, end_bit).
int sort_pairs(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout,
               size_t n, unsigned end_bit, hipStream_t st) {
    size_t tmp_bytes = 0;
    MPK_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, n, 0u,
                                          end_bit, st));
    DevBuf tmp;
    MPK_HIP_TRY(tmp.alloc(tmp_bytes));
    MPK_HIP_TRY(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, kin, kout, vin, vout, n, 0u,
                                          end_bit, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));  // before tmp is freed
    return MPK_OK;
}

template <class T>
int inclusive_scan(const T* in, T* out, size_t n, hipStream_t st) {
    size_t tmp_bytes = 0;
    MPK_HIP_TRY(rocprim::inclusive_scan(nullptr, tmp_bytes, in, out, n, rocprim::plus<T>(), st));
    DevBuf tmp;
    MPK_HIP_TRY(tmp.alloc(tmp_bytes));
    MPK_HIP_TRY(rocprim::inclusive_scan(tmp.p, tmp_bytes, in, out, n, rocprim::plus<T>(), st));
    MPK_HIP_TRY(hipStreamSynchronize(st));  // before tmp is freed
    return MPK_OK;
}

template <class T>
int exclusive_scan(const T* in, T* out, size_t n, hipStream_t st) {
    size_t tmp_bytes = 0;
    MPK_HIP_TRY(rocprim::exclusive_scan(nullptr, tmp_bytes, in, out, T(0), n, rocprim::plus<T>(),
                                        st));
    DevBuf tmp;
    MPK_HIP_TRY(tmp.alloc(tmp_bytes));
    MPK_HIP_TRY(rocprim::exclusive_scan(tmp.p, tmp_bytes, in, out, T(0), n, rocprim::plus<T>(),
                                        st));
    MPK_HIP_TRY(hipStreamSynchronize(st));  // before tmp is freed
    return MPK_OK;
}

template <class T>
int to_host(T* h, const T* d, size_t n, hipStream_t st) {
    MPK_HIP_TRY(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    return MPK_OK;
}

// ---------------------------------------------------------------- kernels --

// Number of keys < limit in a sorted array (single thread, binary search).
__global__ void k_count_below(const uint64_t* __restrict__ keys, int64_t n, uint64_t limit,
                              int64_t* __restrict__ out) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < limit) lo = mid + 1;
        else hi = mid;
    }
    *out = lo;
}

// head[k] = 1 when sorted key k starts a new run (k < nvalid), else 0.
__global__ void k_heads(const uint64_t* __restrict__ keys, const int64_t* __restrict__ nvalid_p,
                        int64_t n, int32_t* __restrict__ head) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const int64_t nvalid = *nvalid_p;
    head[k] = (k < nvalid && (k == 0 || keys[k] != keys[k - 1])) ? 1 : 0;
}

// --- edges ---
__global__ void k_edge_keys(const int32_t* __restrict__ T, int64_t ne, int64_t nn,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                            int8_t* __restrict__ sign, int* __restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= 6 * ne) return;
    const int64_t e = i / 6;
    const int k = (int)(i - 6 * e);
    const int32_t a = T[4 * e + edge_a(k)];
    const int32_t b = T[4 * e + edge_b(k)];
    if (a < 1 || a > nn || b < 1 || b > nn) atomicOr(err, 1);
    const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    keys[i] = (uint64_t)(uint32_t)(lo - 1) * (uint64_t)nn + (uint64_t)(uint32_t)(hi - 1);
    vals[i] = (uint32_t)i;
    sign[i] = (a == lo) ? (int8_t)1 : (int8_t)-1;
}

__global__ void k_edge_scatter(const uint64_t* __restrict__ keys,
                               const uint32_t* __restrict__ vals,
                               const int32_t* __restrict__ head, const int32_t* __restrict__ rank,
                               int64_t n, int64_t nn, int32_t* __restrict__ edges,
                               int32_t* __restrict__ t2e) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const int32_t r = rank[k];  // inclusive scan of heads == 1-based edge id
    t2e[vals[k]] = r;
    if (head[k]) {
        const uint64_t key = keys[k];
        edges[2 * (int64_t)(r - 1) + 0] = (int32_t)(key / (uint64_t)nn) + 1;
        edges[2 * (int64_t)(r - 1) + 1] = (int32_t)(key % (uint64_t)nn) + 1;
    }
}

// --- dofs ---
__global__ void k_mark_active(const int32_t* __restrict__ t2e, int64_t ne, int64_t nedge,
                              const uint8_t* __restrict__ active, int32_t* __restrict__ mark,
                              int* __restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= 6 * ne) return;
    const int64_t e = i / 6;
    if (active && !active[e]) return;
    const int32_t ed = t2e[i];
    if (ed > nedge || ed < 0) {
        atomicOr(err, 1);
        return;
    }
    if (ed > 0) mark[ed - 1] = 1;  // benign same-value race
}

__global__ void k_entity_map(const int32_t* __restrict__ mark, const int32_t* __restrict__ scan,
                             int64_t nedge, int64_t offset, int32_t* __restrict__ emap) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= nedge) return;
    emap[i] = mark[i] ? (int32_t)(offset + scan[i]) : 0;
}

__global__ void k_cell_dofs(const int32_t* __restrict__ t2e, int64_t ne, int64_t nedge,
                            const int32_t* __restrict__ emap, int32_t* __restrict__ cd) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= 6 * ne) return;
    const int32_t ed = t2e[i];
    cd[i] = (ed > 0 && ed <= nedge) ? emap[ed - 1] : 0;
}

// --- sparse() ---
__global__ void k_coo_keys(const double* __restrict__ I, const double* __restrict__ J,
                           int64_t nt, int64_t m, int64_t n, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ vals, int* __restrict__ err) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= nt) return;
    const double i = I[t], j = J[t];
    uint64_t key = (uint64_t)m * (uint64_t)n;  // invalid: sorts after every valid key
    if (i > 0 && j > 0) {
        const int64_t ii = (int64_t)i, jj = (int64_t)j;
        if (ii > m || jj > n || (double)ii != i || (double)jj != j) atomicOr(err, 1);
        else key = (uint64_t)(jj - 1) * (uint64_t)m + (uint64_t)(ii - 1);
    }
    keys[t] = key;
    vals[t] = (uint32_t)t;
}

// One thread per run: sums V in sorted (== triplet) order.
__global__ void k_run_sums(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                           const int64_t* __restrict__ runstart, const int64_t* __restrict__ nrun_p,
                           const int64_t* __restrict__ nvalid_p, const double* __restrict__ V,
                           double* __restrict__ sums, int32_t* __restrict__ keep, int drop) {
    const int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t nrun = *nrun_p;
    if (g >= nrun) return;
    const int64_t b = runstart[g];
    const int64_t e = (g + 1 < nrun) ? runstart[g + 1] : *nvalid_p;
    double s = 0.0;
    for (int64_t k = b; k < e; ++k) s += V[vals[k]];
    sums[g] = s;
    keep[g] = (!drop || s != 0.0) ? 1 : 0;
}

__global__ void k_run_starts(const int32_t* __restrict__ head, const int32_t* __restrict__ seg,
                             int64_t n, int64_t* __restrict__ runstart) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    if (head[k]) runstart[seg[k] - 1] = k;
}

__global__ void k_csc_emit(const uint64_t* __restrict__ keys,
                           const int64_t* __restrict__ runstart, const int64_t* __restrict__ nrun_p,
                           const double* __restrict__ sums, const int32_t* __restrict__ keep,
                           const int32_t* __restrict__ keep_scan, int64_t m,
                           int32_t* __restrict__ rowind, double* __restrict__ out_vals,
                           int64_t* __restrict__ colcount) {
    const int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g >= *nrun_p || !keep[g]) return;
    const uint64_t key = keys[runstart[g]];
    const int64_t p = keep_scan[g] - 1;
    rowind[p] = (int32_t)(key % (uint64_t)m);
    out_vals[p] = sums[g];
    atomicAdd((unsigned long long*)&colcount[key / (uint64_t)m], 1ull);
}

// --- plan ---
__global__ void k_plan_T(const int32_t* __restrict__ T, int64_t ne, int64_t nn,
                         int4* __restrict__ T4, int* __restrict__ err) {
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne) return;
    int4 t = make_int4(T[4 * e] - 1, T[4 * e + 1] - 1, T[4 * e + 2] - 1, T[4 * e + 3] - 1);
    if (t.x < 0 || t.x >= nn || t.y < 0 || t.y >= nn || t.z < 0 || t.z >= nn || t.w < 0 ||
        t.w >= nn) {
        atomicOr(err, 1);
        t = make_int4(0, 0, 0, 0);
    }
    T4[e] = t;
}

// 0-based copy of the cell DOFs (-1: inactive); values outside [0, n] are
// caught by k_plan_keys before any kernel reads this copy.
__global__ void k_plan_cell(const int32_t* __restrict__ CD, int64_t ne, int32_t* __restrict__ cell) {
    const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (s >= 6 * ne) return;
    cell[s] = CD[s] - 1;
}

__global__ void k_plan_keys(const int32_t* __restrict__ CD, int64_t ne, int64_t n,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                            int* __restrict__ err) {
    const int64_t s = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (s >= 36 * ne) return;
    const int64_t e = s / 36;
    const int rc = (int)(s - 36 * e);
    const int c = rc / 6, r = rc - 6 * c;
    const int32_t dr = CD[6 * e + r], dc = CD[6 * e + c];
    uint64_t key = (uint64_t)n * (uint64_t)n;
    if (dr < 0 || dc < 0 || dr > n || dc > n) atomicOr(err, 2);
    else if (dr > 0 && dc > 0) key = (uint64_t)(dc - 1) * (uint64_t)n + (uint64_t)(dr - 1);
    keys[s] = key;
    vals[s] = (uint32_t)s;
}

__global__ void k_plan_rows(const uint64_t* __restrict__ keys, const int32_t* __restrict__ head,
                            const int32_t* __restrict__ seg, int64_t nsorted, int64_t n,
                            int32_t* __restrict__ rowind, int64_t* __restrict__ segstart,
                            int32_t* __restrict__ colcount) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= nsorted || !head[k]) return;
    const uint64_t key = keys[k];
    const int64_t p = seg[k] - 1;
    rowind[p] = (int32_t)(key % (uint64_t)n);
    segstart[p] = k;
    atomicAdd(&colcount[key / (uint64_t)n], 1);
}

// Diagonal triplets (r == c) are the incidences: count them per column.
__global__ void k_plan_inc_count(const uint64_t* __restrict__ keys,
                                 const uint32_t* __restrict__ vals, const int64_t* __restrict__ nvalid_p,
                                 int64_t n, int32_t* __restrict__ inccount) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= *nvalid_p) return;
    const uint32_t s = vals[k];
    const int rc = (int)(s % 36u);
    if (rc % 7 != 0) return;  // r == c  <=>  (6c + r) % 7 == 0 for r, c in [0, 6)
    atomicAdd(&inccount[keys[k] / (uint64_t)n], 1);
}

// Incidence flag per sorted triplet: a diagonal triplet (r == c) of a valid key.
__global__ void k_plan_inc_flags(const uint32_t* __restrict__ vals,
                                 const int64_t* __restrict__ nvalid_p, int64_t nsorted,
                                 int32_t* __restrict__ flag) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= nsorted) return;
    flag[k] = (k < *nvalid_p && (vals[k] % 36u) % 7u == 0u) ? 1 : 0;
}

// Incidence lists.  Sorted order groups keys by column, and every incidence
// of column j sits in the (j, j) run in ascending slot (== element) order,
// so the global incidence index is the running count of incidence flags.
__global__ void k_plan_inc_fill(const uint32_t* __restrict__ vals,
                                const int32_t* __restrict__ flag,
                                const int32_t* __restrict__ flagscan, int64_t nsorted,
                                uint32_t* __restrict__ inc_slot, int32_t* __restrict__ incidx) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= nsorted || !flag[k]) return;
    const uint32_t s = vals[k];
    const uint32_t e = s / 36u;
    const int c = (int)(s - 36u * e) / 6;
    const int64_t idx = flagscan[k] - 1;
    inc_slot[idx] = 6u * e + (uint32_t)c;
    incidx[6 * (int64_t)e + c] = (int32_t)idx;
}

__global__ void k_plan_pos(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                           const int32_t* __restrict__ seg, const int64_t* __restrict__ nvalid_p,
                           int64_t n, const int32_t* __restrict__ colptr,
                           const int32_t* __restrict__ incidx, uint8_t* __restrict__ posb,
                           int* __restrict__ err) {
    const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (k >= *nvalid_p) return;
    const uint32_t s = vals[k];
    const uint32_t e = s / 36u;
    const int rc = (int)(s - 36u * e);
    const int c = rc / 6, r = rc - 6 * c;
    const int64_t col = (int64_t)(keys[k] / (uint64_t)n);
    const int64_t pos = (int64_t)(seg[k] - 1) - colptr[col];
    const int32_t ii = incidx[6 * (int64_t)e + c];
    if (ii < 0 || pos < 0 || pos > 254) {
        atomicOr(err, 4);
        return;
    }
    posb[8 * (int64_t)ii + r] = (uint8_t)pos;
}

// Frame-record slots per tile: 64 x (largest incidence count of its columns).
__global__ void k_frame_tile_size(const int32_t* __restrict__ inc_ptr, int64_t n,
                                  int32_t* __restrict__ size) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t j0 = t * kTileCols;
    if (j0 >= n) return;
    const int64_t j1 = j0 + kTileCols < n ? j0 + kTileCols : n;
    int m = 0;
    for (int64_t j = j0; j < j1; ++j) m = max(m, inc_ptr[j + 1] - inc_ptr[j]);
    size[t] = m * kTileCols;
}

// Frame records of the closed-form kernels (layout in mpk_plan.hpp), one
// thread per column, written k-major into the column's tile.
__global__ void k_plan_frame(const int32_t* __restrict__ inc_ptr,
                             const uint32_t* __restrict__ slot, const uint2* __restrict__ pos,
                             const int4* __restrict__ T, const int32_t* __restrict__ frame_base,
                             int64_t n, uint4* __restrict__ rec) {
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int64_t base = frame_base[j / kTileCols] + (j % kTileCols);
    for (int k = inc_ptr[j]; k < inc_ptr[j + 1]; ++k) {
        const uint32_t s = slot[k];
        const uint32_t e = s / 6u;
        const int c = (int)(s - 6u * e);
        const uint2 pw = pos[k];
        const int4 t = T[e];
        const int nd[4] = {t.x, t.y, t.z, t.w};
        // frame nodes (lo, hi, o1, o2) as local indices
        int fr[4];
        const int a = edge_a(c), b = edge_b(c);
        fr[0] = nd[a] <= nd[b] ? a : b;
        fr[1] = nd[a] <= nd[b] ? b : a;
        fr[2] = perm_node(c, 2);
        fr[3] = perm_node(c, 3);
        uint32_t fmask = 0, p[6];
        for (int rp = 0; rp < 6; ++rp) {
            const int u = fr[edge_a(rp)], v = fr[edge_b(rp)];
            if (nd[u] < nd[v]) fmask |= 1u << rp;
            const int r = u < v ? edge_index(u, v) : edge_index(v, u);
            p[rp] = r < 4 ? ((pw.x >> (8 * r)) & 0xffu) : ((pw.y >> (8 * (r - 4))) & 0xffu);
        }
        rec[base + (int64_t)(k - inc_ptr[j]) * kTileCols] =
            make_uint4(e, (uint32_t)c | (fmask << 8) | (p[0] << 16) | (p[1] << 24),
                       p[2] | (p[3] << 8) | (p[4] << 16) | (p[5] << 24), 0u);
    }
}

// Ring records (layout in mpk_plan.hpp), one thread per column.  The column
// edge (lo, hi) is shared by all its elements; each element adds two
// opposite nodes.  For a manifold mesh the elements around an edge form a
// closed ring (interior edge: every opposite node seen twice) or an open
// chain (boundary edge: the two chain ends seen once); the walk starts at a
// chain end (or at the first element's first opposite node) and follows the
// shared opposite nodes.  Anything else sets err.
constexpr int kRingMax = 32;

__global__ void k_plan_ring(const int32_t* __restrict__ inc_ptr,
                            const uint32_t* __restrict__ slot, const uint2* __restrict__ pos,
                            const int4* __restrict__ T, const int32_t* __restrict__ frame_base,
                            int64_t n, uint4* __restrict__ rec, int2* __restrict__ ends,
                            int* __restrict__ err) {
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int kb = inc_ptr[j], d = inc_ptr[j + 1] - kb;
    ends[j] = make_int2(0, 0);
    if (d == 0) return;
    {
        const uint32_t s = slot[kb];
        const int c = (int)(s % 6u);
        const int4 t = T[s / 6u];
        const int nd[4] = {t.x, t.y, t.z, t.w};
        const int na = nd[edge_a(c)], nb = nd[edge_b(c)];
        ends[j] = na <= nb ? make_int2(na, nb) : make_int2(nb, na);
    }
    if (d > kRingMax) {
        atomicOr(err, 1);
        return;
    }
    int o1[kRingMax], o2[kRingMax];
    bool used[kRingMax];
    for (int i = 0; i < d; ++i) {
        const uint32_t s = slot[kb + i];
        const int c = (int)(s % 6u);
        const int4 t = T[s / 6u];
        const int nd[4] = {t.x, t.y, t.z, t.w};
        o1[i] = nd[perm_node(c, 2)];
        o2[i] = nd[perm_node(c, 3)];
        used[i] = false;
    }
    auto count = [&](int v) {
        int m = 0;
        for (int i = 0; i < d; ++i) m += (o1[i] == v) + (o2[i] == v);
        return m;
    };
    int start = o1[0];
    for (int i = 0; i < d; ++i) {
        const int c1 = count(o1[i]), c2 = count(o2[i]);
        if (c1 > 2 || c2 > 2) {
            atomicOr(err, 1);
            return;
        }
        if (c1 == 1) { start = o1[i]; break; }
        if (c2 == 1) { start = o2[i]; break; }
    }
    const int64_t base = frame_base[j / kTileCols] + (j % kTileCols);
    int cur = start;
    for (int k = 0; k < d; ++k) {
        int i = -1;
        for (int q = 0; q < d; ++q)
            if (!used[q] && (o1[q] == cur || o2[q] == cur)) {
                i = q;
                break;
            }
        if (i < 0) {
            atomicOr(err, 1);
            return;
        }
        used[i] = true;
        const int onew = (o1[i] == cur) ? o2[i] : o1[i];
        const uint32_t s = slot[kb + i];
        const uint32_t e = s / 6u;
        const int c = (int)(s - 6u * e);
        const uint2 pw = pos[kb + i];
        const int4 t = T[e];
        const int nd[4] = {t.x, t.y, t.z, t.w};
        const int a = edge_a(c), b = edge_b(c);
        // frame (lo, hi, o_prev, o_new) as local node indices
        int fr[4];
        fr[0] = nd[a] <= nd[b] ? a : b;
        fr[1] = nd[a] <= nd[b] ? b : a;
        fr[2] = perm_node(c, 2);
        fr[3] = perm_node(c, 3);
        if (nd[fr[2]] != cur) {
            const int tmp = fr[2];
            fr[2] = fr[3];
            fr[3] = tmp;
        }
        uint32_t fmask = 0, p[6];
        for (int rp = 0; rp < 6; ++rp) {
            const int u = fr[edge_a(rp)], v = fr[edge_b(rp)];
            if (nd[u] < nd[v]) fmask |= 1u << rp;
            const int r = u < v ? edge_index(u, v) : edge_index(v, u);
            p[rp] = r < 4 ? ((pw.x >> (8 * r)) & 0xffu) : ((pw.y >> (8 * (r - 4))) & 0xffu);
        }
        rec[base + (int64_t)k * kTileCols] =
            make_uint4(e | (fmask << 24), (uint32_t)onew | (p[0] << 24),
                       p[1] | (p[2] << 8) | (p[3] << 16) | (p[4] << 24),
                       p[5] | ((uint32_t)start << 8));
        cur = onew;
    }
}

// Compact ring records and column heads (layout in mpk_plan.hpp) from the
// 16-byte ring records, one thread per column.  The conversion checks what
// the compact form derives instead of storing: the orientation bits computed
// from the four node ids must equal the stored ones, every record's
// (lo, o_prev) / (hi, o_prev) positions must be the previous record's
// (lo, o_new) / (hi, o_new) ones, and pos'_0 must be the same for the whole
// column.  Any mismatch or field out of range sets err (the plan then keeps
// the 16-byte records).
__global__ void k_plan_ring8(const int32_t* __restrict__ inc_ptr,
                             const int32_t* __restrict__ colptr,
                             const int32_t* __restrict__ frame_base, int64_t n,
                             const uint4* __restrict__ rec16, const int2* __restrict__ ends,
                             uint2* __restrict__ rec8, uint4* __restrict__ head,
                             int* __restrict__ err) {
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int d = inc_ptr[j + 1] - inc_ptr[j];
    const int64_t j0 = (j / kTileCols) * kTileCols;
    const int cofs = colptr[j] - colptr[j0];
    const int2 lh = ends[j];
    if (d > 255 || cofs > 0xffff || lh.x < 0 || lh.y >= (1 << 24)) {
        atomicOr(err, 1);
        return;
    }
    const int64_t base = frame_base[j / kTileCols] + (j % kTileCols);
    auto pos_of = [](uint4 r, int rp) -> uint32_t {
        // ring record: pos'_0 in y's top byte, pos'_1..pos'_4 in z, pos'_5 in w
        return rp == 0 ? (r.y >> 24) : rp < 5 ? ((r.z >> (8 * (rp - 1))) & 0xffu) : (r.w & 0xffu);
    };
    auto six = [&](uint32_t p) -> uint32_t {
        if (p == 0xffu) return 63u;
        if (p > 62u) atomicOr(err, 1);
        return p & 63u;
    };
    uint32_t p0 = 0xffu, p1 = 0xffu, p3 = 0xffu, q2 = 0, q4 = 0;
    int ostart = 0, oprev = 0;
    for (int k = 0; k < d; ++k) {
        const uint4 r = rec16[base + (int64_t)k * kTileCols];
        const uint32_t e = r.x & 0xffffffu;
        const int onew = (int)(r.y & 0xffffffu);
        if (k == 0) {
            p0 = pos_of(r, 0);
            p1 = pos_of(r, 1);
            p3 = pos_of(r, 3);
            ostart = oprev = (int)(r.w >> 8);
        } else if (pos_of(r, 0) != p0 || pos_of(r, 1) != q2 || pos_of(r, 3) != q4) {
            atomicOr(err, 1);
        }
        const uint32_t f = 1u | (uint32_t)(lh.x < oprev) << 1 | (uint32_t)(lh.x < onew) << 2 |
                           (uint32_t)(lh.y < oprev) << 3 | (uint32_t)(lh.y < onew) << 4 |
                           (uint32_t)(oprev < onew) << 5;
        const int dn = onew - lh.x;
        if (f != ((r.x >> 24) & 0x3fu) || dn < -(1 << 21) || dn >= (1 << 21)) atomicOr(err, 1);
        const uint32_t s2 = six(pos_of(r, 2)), s4 = six(pos_of(r, 4)), s5 = six(pos_of(r, 5));
        rec8[base + (int64_t)k * kTileCols] =
            make_uint2(e | s2 << 24 | (s5 & 3u) << 30,
                       ((uint32_t)dn & 0x3fffffu) | s4 << 22 | (s5 >> 2) << 28);
        q2 = pos_of(r, 2);
        q4 = pos_of(r, 4);
        oprev = onew;
    }
    head[j] = make_uint4((uint32_t)lh.x | p0 << 24, (uint32_t)lh.y | p1 << 24,
                         (uint32_t)ostart | p3 << 24, (uint32_t)cofs | (uint32_t)d << 16);
}

__global__ void k_tile_max(const int32_t* __restrict__ colptr, int64_t n,
                           int* __restrict__ maxv) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t j0 = t * kTileCols;
    if (j0 >= n) return;
    const int64_t j1 = j0 + kTileCols < n ? j0 + kTileCols : n;
    atomicMax(maxv, colptr[j1] - colptr[j0]);
}

// --- drop zeros ---
__global__ void k_nonzero_flags(const double* __restrict__ v, int64_t nnz, int cplx,
                                int32_t* __restrict__ f) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= nnz) return;
    f[p] = cplx ? (v[2 * p] != 0.0 || v[2 * p + 1] != 0.0) : (v[p] != 0.0);
}

__global__ void k_compact(const int32_t* __restrict__ f, const int32_t* __restrict__ fs,
                          const int32_t* __restrict__ rowind, const double* __restrict__ v,
                          int64_t nnz, int cplx, int32_t* __restrict__ ro,
                          double* __restrict__ vo) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= nnz || !f[p]) return;
    const int64_t q = fs[p] - 1;
    ro[q] = rowind[p];
    if (cplx) {
        vo[2 * q] = v[2 * p];
        vo[2 * q + 1] = v[2 * p + 1];
    } else {
        vo[q] = v[p];
    }
}

__global__ void k_new_colptr(const int32_t* __restrict__ colptr, const int32_t* __restrict__ fs,
                             int64_t n, int64_t* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j > n) return;
    const int32_t p = colptr[j];
    out[j] = p == 0 ? 0 : fs[p - 1];
}

int check_err(const int* err_dev, hipStream_t st, int* h) {
    *h = 0;
    MPK_HIP_TRY(hipMemcpyAsync(h, err_dev, sizeof(int), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    return MPK_OK;
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_generate_edges(const int32_t* T, int64_t ne, int64_t nn, int32_t* edges_out,
                       int32_t* t2e_out, int8_t* sign_out, int64_t* nedge_out, void* stream) {
    clear_error();
    MPK_REQUIRE(ne >= 0 && nn >= 0 && nedge_out, MPK_ERR_ARGS, "bad size or NULL nedge_out");
    *nedge_out = 0;
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(T && edges_out && t2e_out && sign_out, MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(6 * ne < (int64_t)UINT32_MAX, MPK_ERR_LIMIT, "6*ne exceeds 2^32");
    MPK_REQUIRE(nn < (int64_t)INT32_MAX, MPK_ERR_LIMIT, "nn exceeds 2^31");
    hipStream_t st = as_stream(stream);
    const int64_t nr = 6 * ne;
    DevBuf k0, k1, v0, v1, head, rank, err;
    MPK_HIP_TRY(k0.alloc(nr * 8)); MPK_HIP_TRY(k1.alloc(nr * 8));
    MPK_HIP_TRY(v0.alloc(nr * 4)); MPK_HIP_TRY(v1.alloc(nr * 4));
    MPK_HIP_TRY(head.alloc(nr * 4)); MPK_HIP_TRY(rank.alloc(nr * 4));
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_edge_keys, dim3(grid_for(nr, kB)), dim3(kB), 0, st, T, ne, nn,
                       k0.as<uint64_t>(), v0.as<uint32_t>(), sign_out, err.as<int>());
    MPK_HIP_TRY(hipGetLastError());
    int h = 0;
    if (int rc = check_err(err.as<int>(), st, &h)) return rc;
    MPK_REQUIRE(h == 0, MPK_ERR_INDEX, "node id outside [1, nn]");
    const unsigned nb = bits_for((uint64_t)nn * (uint64_t)nn);
    if (int rc = sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(),
                            v1.as<uint32_t>(), nr, nb, st))
        return rc;
    // all keys are valid: nvalid == nr
    DevBuf nv;
    MPK_HIP_TRY(nv.alloc(sizeof(int64_t)));
    MPK_HIP_TRY(hipMemcpyAsync(nv.p, &nr, sizeof(int64_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_heads, dim3(grid_for(nr, kB)), dim3(kB), 0, st, k1.as<uint64_t>(),
                       nv.as<int64_t>(), nr, head.as<int32_t>());
    if (int rc = inclusive_scan(head.as<int32_t>(), rank.as<int32_t>(), nr, st)) return rc;
    hipLaunchKernelGGL(k_edge_scatter, dim3(grid_for(nr, kB)), dim3(kB), 0, st,
                       k1.as<uint64_t>(), v1.as<uint32_t>(), head.as<int32_t>(),
                       rank.as<int32_t>(), nr, nn, edges_out, t2e_out);
    MPK_HIP_TRY(hipGetLastError());
    int32_t last = 0;
    if (int rc = to_host(&last, rank.as<int32_t>() + nr - 1, 1, st)) return rc;
    *nedge_out = last;
    return MPK_OK;
}

int mpk_distribute_dofs(const int32_t* t2e, int64_t ne, int64_t nedge,
                        const uint8_t* active_elem, int64_t offset, int32_t* entity_map,
                        int32_t* cell_dofs, int64_t* ndofs_out, void* stream) {
    clear_error();
    MPK_REQUIRE(ne >= 0 && nedge >= 0 && offset >= 0 && ndofs_out, MPK_ERR_ARGS, "bad argument");
    *ndofs_out = 0;
    MPK_REQUIRE(ne == 0 || (t2e && cell_dofs), MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(nedge == 0 || entity_map, MPK_ERR_ARGS, "NULL entity_map");
    MPK_REQUIRE(offset + nedge < (int64_t)INT32_MAX, MPK_ERR_LIMIT, "DOF ids exceed 2^31");
    hipStream_t st = as_stream(stream);
    DevBuf mark, scan, err;
    MPK_HIP_TRY(mark.alloc((nedge + 1) * 4));
    MPK_HIP_TRY(scan.alloc((nedge + 1) * 4));
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(mark.p, 0, (nedge + 1) * 4, st));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    if (ne > 0)
        hipLaunchKernelGGL(k_mark_active, dim3(grid_for(6 * ne, kB)), dim3(kB), 0, st, t2e, ne,
                           nedge, active_elem, mark.as<int32_t>(), err.as<int>());
    int h = 0;
    if (int rc = check_err(err.as<int>(), st, &h)) return rc;
    MPK_REQUIRE(h == 0, MPK_ERR_INDEX, "T2E entry outside [0, nedge]");
    int64_t nd = 0;
    if (nedge > 0) {
        if (int rc = inclusive_scan(mark.as<int32_t>(), scan.as<int32_t>(), nedge, st)) return rc;
        hipLaunchKernelGGL(k_entity_map, dim3(grid_for(nedge, kB)), dim3(kB), 0, st,
                           mark.as<int32_t>(), scan.as<int32_t>(), nedge, offset, entity_map);
        int32_t last = 0;
        if (int rc = to_host(&last, scan.as<int32_t>() + nedge - 1, 1, st)) return rc;
        nd = last;
    }
    if (ne > 0)
        hipLaunchKernelGGL(k_cell_dofs, dim3(grid_for(6 * ne, kB)), dim3(kB), 0, st, t2e, ne,
                           nedge, entity_map, cell_dofs);
    MPK_HIP_TRY(hipGetLastError());
    MPK_HIP_TRY(hipStreamSynchronize(st));
    *ndofs_out = nd;
    return MPK_OK;
}

int mpk_sparse_coo_to_csc(const double* I, const double* J, const double* V, int64_t nt,
                          int64_t m, int64_t n, int drop_zeros, int64_t* colptr_out,
                          int32_t* rowind_out, double* vals_out, int64_t cap,
                          int64_t* nnz_out, void* stream) {
    clear_error();
    MPK_REQUIRE(nt >= 0 && m >= 0 && n >= 0 && cap >= 0 && nnz_out && colptr_out, MPK_ERR_ARGS,
                "bad argument");
    MPK_REQUIRE(nt < (int64_t)UINT32_MAX, MPK_ERR_LIMIT, "more than 2^32 triplets");
    MPK_REQUIRE(m < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX, MPK_ERR_LIMIT,
                "matrix dimension exceeds 2^31");
    hipStream_t st = as_stream(stream);
    *nnz_out = 0;
    if (nt == 0 || m == 0 || n == 0) {
        MPK_HIP_TRY(hipMemsetAsync(colptr_out, 0, (n + 1) * sizeof(int64_t), st));
        MPK_HIP_TRY(hipStreamSynchronize(st));
        return MPK_OK;
    }
    MPK_REQUIRE(I && J && V, MPK_ERR_ARGS, "NULL triplet array");
    DevBuf k0, k1, v0, v1, head, seg, err, nv, runstart, sums, keep, keepscan, colcount, nrun;
    MPK_HIP_TRY(k0.alloc(nt * 8)); MPK_HIP_TRY(k1.alloc(nt * 8));
    MPK_HIP_TRY(v0.alloc(nt * 4)); MPK_HIP_TRY(v1.alloc(nt * 4));
    MPK_HIP_TRY(head.alloc(nt * 4)); MPK_HIP_TRY(seg.alloc(nt * 4));
    MPK_HIP_TRY(err.alloc(sizeof(int))); MPK_HIP_TRY(nv.alloc(sizeof(int64_t)));
    MPK_HIP_TRY(nrun.alloc(sizeof(int64_t)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_coo_keys, dim3(grid_for(nt, kB)), dim3(kB), 0, st, I, J, nt, m, n,
                       k0.as<uint64_t>(), v0.as<uint32_t>(), err.as<int>());
    MPK_HIP_TRY(hipGetLastError());
    int h = 0;
    if (int rc = check_err(err.as<int>(), st, &h)) return rc;
    MPK_REQUIRE(h == 0, MPK_ERR_INDEX, "sparse(): index exceeds matrix dimensions or is not an integer");
    const uint64_t inval = (uint64_t)m * (uint64_t)n;
    if (int rc = sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(),
                            v1.as<uint32_t>(), nt, bits_for(inval), st))
        return rc;
    hipLaunchKernelGGL(k_count_below, dim3(1), dim3(1), 0, st, k1.as<uint64_t>(), nt, inval,
                       nv.as<int64_t>());
    hipLaunchKernelGGL(k_heads, dim3(grid_for(nt, kB)), dim3(kB), 0, st, k1.as<uint64_t>(),
                       nv.as<int64_t>(), nt, head.as<int32_t>());
    if (int rc = inclusive_scan(head.as<int32_t>(), seg.as<int32_t>(), nt, st)) return rc;
    int32_t nruns = 0;
    if (int rc = to_host(&nruns, seg.as<int32_t>() + nt - 1, 1, st)) return rc;
    const int64_t nr = nruns;
    MPK_HIP_TRY(hipMemcpyAsync(nrun.p, &nr, sizeof(int64_t), hipMemcpyHostToDevice, st));
    MPK_HIP_TRY(runstart.alloc((nr + 1) * 8)); MPK_HIP_TRY(sums.alloc((nr + 1) * 8));
    MPK_HIP_TRY(keep.alloc((nr + 1) * 4)); MPK_HIP_TRY(keepscan.alloc((nr + 1) * 4));
    MPK_HIP_TRY(colcount.alloc((n + 1) * 8));
    MPK_HIP_TRY(hipMemsetAsync(colcount.p, 0, (n + 1) * 8, st));
    int64_t nnz = 0;
    if (nr > 0) {
        hipLaunchKernelGGL(k_run_starts, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           head.as<int32_t>(), seg.as<int32_t>(), nt, runstart.as<int64_t>());
        hipLaunchKernelGGL(k_run_sums, dim3(grid_for(nr, kB)), dim3(kB), 0, st,
                           k1.as<uint64_t>(), v1.as<uint32_t>(), runstart.as<int64_t>(),
                           nrun.as<int64_t>(), nv.as<int64_t>(), V, sums.as<double>(),
                           keep.as<int32_t>(), drop_zeros);
        if (int rc = inclusive_scan(keep.as<int32_t>(), keepscan.as<int32_t>(), nr, st)) return rc;
        int32_t kept = 0;
        if (int rc = to_host(&kept, keepscan.as<int32_t>() + nr - 1, 1, st)) return rc;
        nnz = kept;
    }
    *nnz_out = nnz;
    if (nnz > cap) {
        set_error("sparse(): nnz=%lld exceeds capacity %lld", (long long)nnz, (long long)cap);
        return MPK_ERR_LIMIT;
    }
    if (nr > 0) {
        hipLaunchKernelGGL(k_csc_emit, dim3(grid_for(nr, kB)), dim3(kB), 0, st,
                           k1.as<uint64_t>(), runstart.as<int64_t>(), nrun.as<int64_t>(),
                           sums.as<double>(), keep.as<int32_t>(), keepscan.as<int32_t>(), m,
                           rowind_out, vals_out, colcount.as<int64_t>());
    }
    if (int rc = exclusive_scan(colcount.as<int64_t>(), colptr_out, n + 1, st)) return rc;
    MPK_HIP_TRY(hipGetLastError());
    MPK_HIP_TRY(hipStreamSynchronize(st));
    return MPK_OK;
}

int mpk_plan_create(const int32_t* T, const int32_t* cell_dofs, int64_t ne, int64_t nn,
                    int64_t ndofs, void* stream, mpk_plan** plan_out) {
    clear_error();
    MPK_REQUIRE(plan_out, MPK_ERR_ARGS, "NULL plan_out");
    *plan_out = nullptr;
    MPK_REQUIRE(ne >= 0 && nn >= 0 && ndofs >= 0, MPK_ERR_ARGS, "negative size");
    MPK_REQUIRE(ne == 0 || (T && cell_dofs), MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(36 * ne < (int64_t)UINT32_MAX, MPK_ERR_LIMIT, "36*ne exceeds 2^32");
    MPK_REQUIRE(ndofs < (int64_t)INT32_MAX && nn < (int64_t)INT32_MAX, MPK_ERR_LIMIT,
                "more than 2^31 DOFs or nodes");
    hipStream_t st = as_stream(stream);
    auto plan = new mpk_plan();
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(st);  // no kernel may still use plan memory
        delete plan;
        return rc;
    };
    plan->ne = ne; plan->nn = nn; plan->n = ndofs;
    MPK_HIP_TRY(hipGetDevice(&plan->device));
    const int64_t nt = 36 * ne;
    if (plan->buf_T.alloc(ne * sizeof(int4)) != hipSuccess ||
        plan->buf_colptr.alloc((ndofs + 1) * 4) != hipSuccess ||
        plan->buf_incptr.alloc((ndofs + 1) * 4) != hipSuccess) {
        set_error("hipMalloc failed for plan arrays");
        return fail(MPK_ERR_HIP);
    }
    plan->T = plan->buf_T.as<int4>();
    plan->colptr = plan->buf_colptr.as<int32_t>();
    plan->inc_ptr = plan->buf_incptr.as<int32_t>();
    DevBuf err;
    if (err.alloc(sizeof(int)) != hipSuccess) return fail(MPK_ERR_HIP);
    (void)hipMemsetAsync(err.p, 0, sizeof(int), st);
    (void)hipMemsetAsync(plan->colptr, 0, (ndofs + 1) * 4, st);
    (void)hipMemsetAsync(plan->inc_ptr, 0, (ndofs + 1) * 4, st);
    if (ne == 0 || ndofs == 0) {
        (void)hipStreamSynchronize(st);
        *plan_out = plan;
        return MPK_OK;
    }
    hipLaunchKernelGGL(k_plan_T, dim3(grid_for(ne, kB)), dim3(kB), 0, st, T, ne, nn, plan->T,
                       err.as<int>());
    if (plan->buf_cell.alloc(6 * ne * sizeof(int32_t)) != hipSuccess) {
        set_error("hipMalloc failed for the plan's cell DOFs");
        return fail(MPK_ERR_HIP);
    }
    plan->cell = plan->buf_cell.as<int32_t>();
    hipLaunchKernelGGL(k_plan_cell, dim3(grid_for(6 * ne, kB)), dim3(kB), 0, st, cell_dofs, ne,
                       plan->cell);
    {
        DevBuf k0, k1, v0, v1, head, seg, nv, segstart, colcount, inccount, incidx, posb;
        auto A = [&](DevBuf& b, size_t bytes) { return b.alloc(bytes) == hipSuccess; };
        if (!A(k0, nt * 8) || !A(k1, nt * 8) || !A(v0, nt * 4) || !A(v1, nt * 4) ||
            !A(nv, sizeof(int64_t)) || !A(colcount, (ndofs + 1) * 4) ||
            !A(inccount, (ndofs + 1) * 4)) {
            set_error("hipMalloc failed for plan scratch (%lld triplets)", (long long)nt);
            return fail(MPK_ERR_HIP);
        }
        hipLaunchKernelGGL(k_plan_keys, dim3(grid_for(nt, kB)), dim3(kB), 0, st, cell_dofs, ne,
                           ndofs, k0.as<uint64_t>(), v0.as<uint32_t>(), err.as<int>());
        int h = 0;
        if (int rc = check_err(err.as<int>(), st, &h)) return fail(rc);
        if (h & 1) { set_error("node id outside [1, nn]"); return fail(MPK_ERR_INDEX); }
        if (h & 2) { set_error("cell DOF outside [0, ndofs]"); return fail(MPK_ERR_INDEX); }
        const uint64_t inval = (uint64_t)ndofs * (uint64_t)ndofs;
        if (int rc = sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(),
                                v1.as<uint32_t>(), nt, bits_for(inval), st))
            return fail(rc);
        (void)hipStreamSynchronize(st);
        k0.release();
        v0.release();
        if (!A(head, nt * 4) || !A(seg, nt * 4)) { set_error("hipMalloc failed"); return fail(MPK_ERR_HIP); }
        hipLaunchKernelGGL(k_count_below, dim3(1), dim3(1), 0, st, k1.as<uint64_t>(), nt, inval,
                           nv.as<int64_t>());
        hipLaunchKernelGGL(k_heads, dim3(grid_for(nt, kB)), dim3(kB), 0, st, k1.as<uint64_t>(),
                           nv.as<int64_t>(), nt, head.as<int32_t>());
        if (int rc = inclusive_scan(head.as<int32_t>(), seg.as<int32_t>(), nt, st)) return fail(rc);
        int32_t nnz32 = 0;
        if (int rc = to_host(&nnz32, seg.as<int32_t>() + nt - 1, 1, st)) return fail(rc);
        plan->nnz = nnz32;
        if (!A(plan->buf_rowind, (size_t)(plan->nnz + 1) * 4) ||
            !A(segstart, (size_t)(plan->nnz + 1) * 8)) {
            set_error("hipMalloc failed for pattern");
            return fail(MPK_ERR_HIP);
        }
        plan->rowind = plan->buf_rowind.as<int32_t>();
        (void)hipMemsetAsync(colcount.p, 0, (ndofs + 1) * 4, st);
        (void)hipMemsetAsync(inccount.p, 0, (ndofs + 1) * 4, st);
        hipLaunchKernelGGL(k_plan_rows, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           k1.as<uint64_t>(), head.as<int32_t>(), seg.as<int32_t>(), nt, ndofs,
                           plan->rowind, segstart.as<int64_t>(), colcount.as<int32_t>());
        if (int rc = exclusive_scan(colcount.as<int32_t>(), plan->colptr, ndofs + 1, st))
            return fail(rc);
        hipLaunchKernelGGL(k_plan_inc_count, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           k1.as<uint64_t>(), v1.as<uint32_t>(), nv.as<int64_t>(), ndofs,
                           inccount.as<int32_t>());
        if (int rc = exclusive_scan(inccount.as<int32_t>(), plan->inc_ptr, ndofs + 1, st))
            return fail(rc);
        int32_t ninc32 = 0;
        if (int rc = to_host(&ninc32, plan->inc_ptr + ndofs, 1, st)) return fail(rc);
        plan->ninc = ninc32;
        if (!A(plan->buf_incslot, (size_t)(plan->ninc + 1) * 4) ||
            !A(plan->buf_incpos, (size_t)(plan->ninc + 1) * 8) || !A(incidx, (size_t)6 * ne * 4)) {
            set_error("hipMalloc failed for incidences");
            return fail(MPK_ERR_HIP);
        }
        plan->inc_slot = plan->buf_incslot.as<uint32_t>();
        plan->inc_pos = plan->buf_incpos.as<uint2>();
        (void)hipMemsetAsync(incidx.p, 0xff, (size_t)6 * ne * 4, st);
        (void)hipMemsetAsync(plan->inc_pos, 0xff, (size_t)(plan->ninc + 1) * 8, st);
        // head is free now: reuse it for the incidence flags
        DevBuf flagscan;
        if (!A(flagscan, nt * 4)) { set_error("hipMalloc failed"); return fail(MPK_ERR_HIP); }
        hipLaunchKernelGGL(k_plan_inc_flags, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           v1.as<uint32_t>(), nv.as<int64_t>(), nt, head.as<int32_t>());
        if (int rc = inclusive_scan(head.as<int32_t>(), flagscan.as<int32_t>(), nt, st))
            return fail(rc);
        hipLaunchKernelGGL(k_plan_inc_fill, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           v1.as<uint32_t>(), head.as<int32_t>(), flagscan.as<int32_t>(), nt,
                           plan->inc_slot, incidx.as<int32_t>());
        (void)hipStreamSynchronize(st);
        flagscan.release();
        hipLaunchKernelGGL(k_plan_pos, dim3(grid_for(nt, kB)), dim3(kB), 0, st,
                           k1.as<uint64_t>(), v1.as<uint32_t>(), seg.as<int32_t>(),
                           nv.as<int64_t>(), ndofs, plan->colptr, incidx.as<int32_t>(),
                           reinterpret_cast<uint8_t*>(plan->inc_pos), err.as<int>());
        DevBuf tmax;
        if (!A(tmax, sizeof(int))) return fail(MPK_ERR_HIP);
        (void)hipMemsetAsync(tmax.p, 0, sizeof(int), st);
        const int64_t ntiles = (ndofs + kTileCols - 1) / kTileCols;
        hipLaunchKernelGGL(k_tile_max, dim3(grid_for(ntiles, kB)), dim3(kB), 0, st,
                           plan->colptr, ndofs, tmax.as<int>());
        if (hipGetLastError() != hipSuccess) { set_error("plan kernel launch failed"); return fail(MPK_ERR_HIP); }
        int tm = 0;
        if (int rc = to_host(&tm, tmax.as<int>(), 1, st)) return fail(rc);
        plan->max_tile_nnz = tm;
        if (int rc = check_err(err.as<int>(), st, &h)) return fail(rc);
        if (h & 4) {
            set_error("a column holds more than 255 entries (engine limit)");
            return fail(MPK_ERR_LIMIT);
        }
    }
    if (plan->ninc > 0) {
        const int64_t ntiles = (ndofs + kTileCols - 1) / kTileCols;
        DevBuf tsize;
        if (tsize.alloc((ntiles + 1) * 4) != hipSuccess ||
            plan->buf_framebase.alloc((ntiles + 1) * 4) != hipSuccess) {
            set_error("hipMalloc failed for frame tiles");
            return fail(MPK_ERR_HIP);
        }
        (void)hipMemsetAsync(tsize.p, 0, (ntiles + 1) * 4, st);
        hipLaunchKernelGGL(k_frame_tile_size, dim3(grid_for(ntiles, kB)), dim3(kB), 0, st,
                           plan->inc_ptr, ndofs, tsize.as<int32_t>());
        plan->frame_base = plan->buf_framebase.as<int32_t>();
        if (int rc = exclusive_scan(tsize.as<int32_t>(), plan->frame_base, ntiles + 1, st))
            return fail(rc);
        int32_t nslots = 0;
        if (int rc = to_host(&nslots, plan->frame_base + ntiles, 1, st)) return fail(rc);
        if (plan->buf_incframe.alloc((size_t)nslots * 16) != hipSuccess) {
            set_error("hipMalloc failed for frame records");
            return fail(MPK_ERR_HIP);
        }
        plan->inc_frame = plan->buf_incframe.as<uint4>();
        hipLaunchKernelGGL(k_plan_frame, dim3(grid_for(ndofs, kB)), dim3(kB), 0, st,
                           plan->inc_ptr, plan->inc_slot, plan->inc_pos, plan->T,
                           plan->frame_base, ndofs, plan->inc_frame);
        // ring records: ids are packed in 24 bits
        if (ne < (int64_t(1) << 24) && nn < (int64_t(1) << 24)) {
            if (plan->buf_incring.alloc((size_t)nslots * 16) != hipSuccess ||
                plan->buf_colends.alloc((size_t)ndofs * 8) != hipSuccess) {
                set_error("hipMalloc failed for ring records");
                return fail(MPK_ERR_HIP);
            }
            (void)hipMemsetAsync(err.p, 0, sizeof(int), st);
            hipLaunchKernelGGL(k_plan_ring, dim3(grid_for(ndofs, kB)), dim3(kB), 0, st,
                               plan->inc_ptr, plan->inc_slot, plan->inc_pos, plan->T,
                               plan->frame_base, ndofs, plan->buf_incring.as<uint4>(),
                               plan->buf_colends.as<int2>(), err.as<int>());
            int hr = 0;
            if (int rc = check_err(err.as<int>(), st, &hr)) return fail(rc);
            plan->ring = (hr == 0);
            if (plan->ring) {
                plan->inc_ring = plan->buf_incring.as<uint4>();
                plan->col_ends = plan->buf_colends.as<int2>();
            } else {
                plan->buf_incring.release();
                plan->buf_colends.release();
            }
            // compact ring records, unless MPK_RING_FORMAT=16 keeps the
            // 16-byte ones (for comparisons and timing)
            const char* fmt = std::getenv("MPK_RING_FORMAT");
            if (plan->ring && !(fmt && std::strcmp(fmt, "16") == 0)) {
                if (plan->buf_incring8.alloc((size_t)nslots * 8) != hipSuccess ||
                    plan->buf_colhead.alloc((size_t)ndofs * 16) != hipSuccess) {
                    set_error("hipMalloc failed for compact ring records");
                    return fail(MPK_ERR_HIP);
                }
                (void)hipMemsetAsync(err.p, 0, sizeof(int), st);
                hipLaunchKernelGGL(k_plan_ring8, dim3(grid_for(ndofs, kB)), dim3(kB), 0, st,
                                   plan->inc_ptr, plan->colptr, plan->frame_base, ndofs,
                                   plan->inc_ring, plan->col_ends,
                                   plan->buf_incring8.as<uint2>(), plan->buf_colhead.as<uint4>(),
                                   err.as<int>());
                int h8 = 0;
                if (int rc = check_err(err.as<int>(), st, &h8)) return fail(rc);
                plan->ring8 = (h8 == 0);
                if (plan->ring8) {
                    plan->inc_ring8 = plan->buf_incring8.as<uint2>();
                    plan->col_head = plan->buf_colhead.as<uint4>();
                    plan->inc_ring = nullptr;
                    plan->col_ends = nullptr;
                    plan->buf_incring.release();
                    plan->buf_colends.release();
                } else {
                    plan->buf_incring8.release();
                    plan->buf_colhead.release();
                }
            }
        }
    }
    if (plan->buf_P4.alloc((size_t)(nn > 0 ? nn : 1) * 32) != hipSuccess) {
        set_error("hipMalloc failed for padded coordinates");
        return fail(MPK_ERR_HIP);
    }
    plan->P4 = plan->buf_P4.as<double4>();
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        set_error("plan frame kernel failed");
        return fail(MPK_ERR_HIP);
    }
    *plan_out = plan;
    return MPK_OK;
}

int mpk_plan_destroy(mpk_plan* plan) {
    delete plan;
    return MPK_OK;
}

int mpk_plan_info(const mpk_plan* plan, int64_t* nnz, int64_t* ncols, int64_t* ne,
                  int64_t* n_incidences, int64_t* max_tile_nnz) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    if (nnz) *nnz = plan->nnz;
    if (ncols) *ncols = plan->n;
    if (ne) *ne = plan->ne;
    if (n_incidences) *n_incidences = plan->ninc;
    if (max_tile_nnz) *max_tile_nnz = plan->max_tile_nnz;
    return MPK_OK;
}

int mpk_plan_ring_order(const mpk_plan* plan, int* ring) {
    clear_error();
    MPK_REQUIRE(plan && ring, MPK_ERR_ARGS, "NULL argument");
    *ring = plan->ring ? (plan->ring8 ? 2 : 1) : 0;
    return MPK_OK;
}

int mpk_plan_pattern(const mpk_plan* plan, const int32_t** colptr, const int32_t** rowind) {
    clear_error();
    MPK_REQUIRE(plan && colptr && rowind, MPK_ERR_ARGS, "NULL argument");
    *colptr = plan->colptr;
    *rowind = plan->rowind;
    return MPK_OK;
}

int mpk_csc_drop_zeros(const mpk_plan* plan, const double* vals, int complex_vals,
                       int64_t* colptr_out, int32_t* rowind_out, double* vals_out,
                       int64_t* nnz_out, void* stream) {
    clear_error();
    MPK_REQUIRE(plan && colptr_out && nnz_out, MPK_ERR_ARGS, "NULL argument");
    hipStream_t st = as_stream(stream);
    const int64_t nnz = plan->nnz, n = plan->n;
    *nnz_out = 0;
    if (nnz == 0) {
        MPK_HIP_TRY(hipMemsetAsync(colptr_out, 0, (n + 1) * sizeof(int64_t), st));
        MPK_HIP_TRY(hipStreamSynchronize(st));
        return MPK_OK;
    }
    MPK_REQUIRE(vals && rowind_out && vals_out, MPK_ERR_ARGS, "NULL array argument");
    DevBuf f, fs;
    MPK_HIP_TRY(f.alloc(nnz * 4));
    MPK_HIP_TRY(fs.alloc(nnz * 4));
    hipLaunchKernelGGL(k_nonzero_flags, dim3(grid_for(nnz, kB)), dim3(kB), 0, st, vals, nnz,
                       complex_vals, f.as<int32_t>());
    if (int rc = inclusive_scan(f.as<int32_t>(), fs.as<int32_t>(), nnz, st)) return rc;
    hipLaunchKernelGGL(k_compact, dim3(grid_for(nnz, kB)), dim3(kB), 0, st, f.as<int32_t>(),
                       fs.as<int32_t>(), plan->rowind, vals, nnz, complex_vals, rowind_out,
                       vals_out);
    hipLaunchKernelGGL(k_new_colptr, dim3(grid_for(n + 1, kB)), dim3(kB), 0, st, plan->colptr,
                       fs.as<int32_t>(), n, colptr_out);
    MPK_HIP_TRY(hipGetLastError());
    int32_t kept = 0;
    if (int rc = to_host(&kept, fs.as<int32_t>() + nnz - 1, 1, st)) return rc;
    *nnz_out = kept;
    return MPK_OK;
}

}  // extern "C"

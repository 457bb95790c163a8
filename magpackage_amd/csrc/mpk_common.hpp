// Shared helpers of the MI355X assembly engine: status/error plumbing,
// element geometry constants and the device-side B-H material evaluation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "mpk_assembly.h"

namespace mpk {

void set_error(const char* fmt, ...);
void clear_error();

#define MPK_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _mpk_e = (expr);                                                    \
        if (_mpk_e != hipSuccess) {                                                    \
            ::mpk::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_mpk_e), \
                             __FILE__, __LINE__);                                      \
            return MPK_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

#define MPK_REQUIRE(cond, code, ...)        \
    do {                                    \
        if (!(cond)) {                      \
            ::mpk::set_error(__VA_ARGS__);  \
            return (code);                  \
        }                                   \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
    return static_cast<unsigned>((n + block - 1) / block);
}

// Local edge (a, b) node pairs, src/Mesh.m:118 order [12 13 14 23 24 34].
__host__ __device__ constexpr int edge_a(int k) { return k < 3 ? 0 : (k < 5 ? 1 : 2); }
__host__ __device__ constexpr int edge_b(int k) {
    return k == 0 ? 1 : (k == 1 ? 2 : (k == 2 ? 3 : (k == 3 ? 2 : 3)));
}

// Index of local edge (a, b), a < b.
__host__ __device__ constexpr int edge_index(int a, int b) {
    return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}
// Node permutation that makes local edge c the edge (0,1) of a new frame,
// the other two nodes in ascending order:
//   c=0 (0,1,2,3)  c=1 (0,2,1,3)  c=2 (0,3,1,2)  c=3 (1,2,0,3)  c=4 (1,3,0,2)  c=5 (2,3,0,1)
__host__ __device__ constexpr int perm_node(int c, int i) {
    return (int)(((c == 0 ? 0x3210u : c == 1 ? 0x3120u : c == 2 ? 0x2130u
                 : c == 3 ? 0x3021u : c == 4 ? 0x2031u : 0x1032u) >> (4 * i)) & 0xfu);
}
// Original local edge of frame edge rp.
__host__ __device__ constexpr int orig_edge(int c, int rp) {
    const int a = perm_node(c, edge_a(rp)), b = perm_node(c, edge_b(rp));
    return a < b ? edge_index(a, b) : edge_index(b, a);
}
static_assert(orig_edge(0, 0) == 0 && orig_edge(5, 0) == 5 && orig_edge(5, 5) == 0 &&
                  orig_edge(2, 5) == 3 && orig_edge(4, 3) == 2,
              "permutation tables");

// RAII device buffer for library-internal scratch.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t n) {
        release();
        bytes = n;
        return hipMalloc(&p, n > 0 ? n : 1);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// A library-owned second stream with a fork and a join event, for calls that
// run two independent kernels side by side: the caller's stream records
// `fork`, the side stream waits on it, runs its kernel and records `join`, and
// the caller's stream waits on `join` before anything after the call.  The
// stream has the default priority.  Created on first use on the current
// device; calls that use one object must not overlap.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    SideStream() = default;
    SideStream(const SideStream&) = delete;
    SideStream& operator=(const SideStream&) = delete;
    ~SideStream() {
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (s) (void)hipStreamDestroy(s);
    }
    hipError_t ensure() {
        hipError_t e = hipSuccess;
        if (!s) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess && !fork) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        if (e == hipSuccess && !join) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
        return e;
    }
};

// Device view of the packed material tables (MexMaterialUtils.hpp:9-18).
struct MaterialView {
    const double* lin;
    const double* isnon;
    const double* maxb;
    const double* brk;
    const double* coe;
    const double* start;
    const double* cnt;
    const double* coestart;
};

// evaluate_nu_derivative, src/mex/MexMaterialUtils.hpp:24-87: same branch
// structure and operation order.  `brk` and `coe` may point into LDS.
template <class BrkPtr, class CoePtr>
__device__ __forceinline__ void eval_nu(double s, int tag, const MaterialView& m, BrkPtr brk,
                                        CoePtr coe, double& nu, double& dnu) {
#pragma clang fp contract(off)
    const int idx = tag - 1;
    if (m.isnon[idx] == 0.0) {
        nu = m.lin[idx];
        dnu = 0.0;
        return;
    }
    const double max_s = m.maxb[idx];
    const bool is_extrap = (s > max_s);
    const double s_eval = is_extrap ? max_s : s;
    const int start = (int)m.start[idx];
    const int coe_start = (int)m.coestart[idx];
    const int count = (int)m.cnt[idx];
    int k = 0, lo = 0, hi = count - 1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (s_eval >= brk[start + mid]) {
            k = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    if (k >= count) k = count - 1;
    if (k < 0) k = 0;
    const double dx = s_eval - brk[start + k];
    const double c0 = coe[coe_start + k * 4 + 0];
    const double c1 = coe[coe_start + k * 4 + 1];
    const double c2 = coe[coe_start + k * 4 + 2];
    const double c3 = coe[coe_start + k * 4 + 3];
    const double nu_eval = ((c0 * dx + c1) * dx + c2) * dx + c3;
    const double dnu_eval = (3.0 * c0 * dx + 2.0 * c1) * dx + c2;
    double n, d;
    if (is_extrap) {
        n = nu_eval + dnu_eval * (s - max_s);
        d = dnu_eval;
    } else {
        n = nu_eval;
        d = dnu_eval;
    }
    const double nu0 = m.lin[idx];
    if (n > nu0) {
        n = nu0;
        d = 0.0;
    } else if (n < 1e-7) {
        n = nu0;
        d = 0.0;
    }
    nu = n;
    dnu = d;
}

}  // namespace mpk

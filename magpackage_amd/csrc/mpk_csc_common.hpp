// Shared pieces of the column-owner CSC assembly kernels: kernel argument
// structs, small vector math and the XCD-aware tile mapping.
#pragma once

#include <cstring>
#include <type_traits>

#include <cmath>

#include "mpk_plan.hpp"

namespace mpk {
namespace csc {

constexpr int kW = kTileCols;  // lanes per block == columns per tile
constexpr int kMaxLdsBytes = 64 * 1024;
constexpr int kMatLds = 2048;  // B-H table doubles staged in LDS

struct CscQuad {
    int nq;
    double qw[8];
    double ref[144];
    double wsum;    // sum of weights
    double qd, qo;  // symmetric barycentric moments (diag / off-diag mean)
};

struct TileCtx {
    const int32_t* __restrict__ colptr;
    const int32_t* __restrict__ inc_ptr;
    const uint32_t* __restrict__ inc_slot;
    const uint2* __restrict__ inc_pos;
    const uint4* __restrict__ inc_frame;
    const uint4* __restrict__ inc_ring;
    const int32_t* __restrict__ frame_base;
    const int4* __restrict__ T;
    const double4* __restrict__ P4;
    const int32_t* __restrict__ rowind;
    int64_t n;
    const uint2* __restrict__ inc_ring8;  // compact ring records (plan.ring8)
    const uint4* __restrict__ col_head;   // their column heads
};

__device__ __forceinline__ int pos_of(uint2 pw, int r) {
    return r < 4 ? (int)((pw.x >> (8 * r)) & 0xffu) : (int)((pw.y >> (8 * (r - 4))) & 0xffu);
}

template <class T>
__device__ __forceinline__ T sel6(int c, T v0, T v1, T v2, T v3, T v4, T v5) {
    return c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : c == 3 ? v3 : c == 4 ? v4 : v5;
}

// FMA contraction within one expression only (see mpk_csc_frame.hpp).
#pragma clang fp contract(on)
struct Vec3 {
    double x, y, z;
};
__device__ __forceinline__ Vec3 sub(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
struct Vec3f {
    float x, y, z;
};
__device__ __forceinline__ Vec3f tof(Vec3 v) { return Vec3f{(float)v.x, (float)v.y, (float)v.z}; }
__device__ __forceinline__ float dotf(Vec3f a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Kernel-argument copy of a quadrature table plus the rule's weight sum and
// the symmetric barycentric moments the closed-form mass uses:
// Q_xy = sum_q w_q l_x l_y, qd = mean diagonal, qo = mean off-diagonal.
inline int make_csc_quad(const mpk_quadrature* q, bool curl, CscQuad& a) {
    MPK_REQUIRE(q != nullptr, MPK_ERR_ARGS, "quadrature table is NULL");
    MPK_REQUIRE(q->nq >= 1 && q->nq <= 8, MPK_ERR_ARGS, "nq=%d outside [1, 8]", q->nq);
    std::memset(&a, 0, sizeof(a));
    a.nq = q->nq;
    double ws = 0.0;
    for (int i = 0; i < q->nq; ++i) {
        a.qw[i] = q->qw[i];
        ws += q->qw[i];
    }
    a.wsum = ws;
    const double* src = curl ? q->ref_curl : q->ref_val;
    for (int i = 0; i < 18 * q->nq; ++i) a.ref[i] = src[i];
    // barycentric moments Q_xy = sum_q w_q l_x l_y, l_1 = 1 - xi - eta - zeta
    double Q[4][4] = {};
    for (int k = 0; k < q->nq; ++k) {
        const double xi = q->q_pts[3 * k], eta = q->q_pts[3 * k + 1], ze = q->q_pts[3 * k + 2];
        const double l[4] = {((1.0 - xi) - eta) - ze, xi, eta, ze};
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) Q[x][y] += q->qw[k] * l[x] * l[y];
    }
    double qd = 0.0, qo = 0.0;
    for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) (x == y ? qd : qo) += Q[x][y];
    a.qd = qd / 4.0;
    a.qo = qo / 12.0;
    return MPK_OK;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so that each XCD sweeps one contiguous range of column tiles: the
// elements and nodes shared by neighbouring tiles then stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    constexpr int64_t kXcd = 8;
    const int64_t q = nb / kXcd, r = nb % kXcd;
    const int64_t x = b % kXcd, i = b / kXcd;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Accumulator planes of an op: ops with several outputs may keep them as
// separate LDS planes (struct-of-arrays, plane k of tile entry i at
// acc[i + k * size]) by declaring kPlanes; the block's plane stride is then
// set in op.stride.  kPlanes = 0: the op writes no matrix values (no LDS
// accumulator, nothing stored per tile).  Default: one plane of Op::Acc.
template <class Op, class = void>
struct Planes {
    static constexpr int value = 1;
};
template <class Op>
struct Planes<Op, std::void_t<decltype(Op::kPlanes)>> {
    static constexpr int value = Op::kPlanes;
};

// Tile prologue shared by the drivers: zeroes the block's LDS accumulator
// range [colptr[j0], colptr[jend]) (every plane) and runs the op's per-block
// set-up.
struct TileRange {
    int64_t tile_id, j0, jend;
    int base, size;
};
template <class Op, class Acc>
__device__ __forceinline__ void tile_zero(Op& op, Acc* acc, unsigned char* smem,
                                          const TileRange& t) {
    if constexpr (Planes<Op>::value > 1) op.stride = t.size;
    op.block_init(smem + (size_t)op.acc_bytes_max);
    for (int i = threadIdx.x; i < t.size * Planes<Op>::value; i += kW) acc[i] = Op::zero();
    if constexpr (Planes<Op>::value > 0) __syncthreads();
}
template <class Op, class Acc>
__device__ __forceinline__ TileRange tile_begin_at(const TileCtx& tc, Op& op, Acc* acc,
                                                   unsigned char* smem, int64_t tile_id) {
    TileRange t;
    t.tile_id = tile_id;
    t.j0 = t.tile_id * kW;
    t.jend = t.j0 + kW < tc.n ? t.j0 + kW : tc.n;
    t.base = tc.colptr[t.j0];
    t.size = tc.colptr[t.jend] - t.base;
    tile_zero(op, acc, smem, t);
    return t;
}
template <class Op, class Acc>
__device__ __forceinline__ TileRange tile_begin(const TileCtx& tc, Op& op, Acc* acc,
                                                unsigned char* smem) {
    return tile_begin_at(tc, op, acc, smem, xcd_tile(blockIdx.x, gridDim.x));
}

// Output values are written once and never read back by the kernel: the
// stores carry the non-temporal hint, so the output stream passes the caches
// that hold the gathered coordinates and element data.
#ifndef MPK_NT_STORES
#define MPK_NT_STORES 1
#endif
__device__ __forceinline__ void st_out(double* p, double v) {
#if MPK_NT_STORES
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
__device__ __forceinline__ void st_out(double2* p, double2 v) {
#if MPK_NT_STORES
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<d2v*>(p));
#else
    *p = v;
#endif
}

// Writes tile entries [0, size) of an 8-byte output array at out + base,
// two entries per lane and store (16-byte stores) from the first 16-byte
// aligned address on; val(i) is tile entry i.
template <class F>
__device__ __forceinline__ void store_run(double* __restrict__ out, int64_t base, int size,
                                          F val) {
    if (size <= 0) return;
    double* o = out + base;
    const int a = (int)((reinterpret_cast<uintptr_t>(o) >> 3) & 1u);  // entries before
    const int npair = (size - a) >> 1;
    double2* o2 = reinterpret_cast<double2*>(o + a);
    for (int q = threadIdx.x; q < npair; q += kW) {
        const int i = a + 2 * q;
        st_out(o2 + q, make_double2(val(i), val(i + 1)));
    }
    if (threadIdx.x == 0) {
        if (a) st_out(o, val(0));
        if ((size - a) & 1) st_out(o + size - 1, val(size - 1));
    }
}

// Ops may store a finished tile themselves (store_tile(base, acc, size)),
// e.g. to order the stores of several output arrays.
template <class Op, class = void>
struct HasStoreTile : std::false_type {};
template <class Op>
struct HasStoreTile<Op, std::void_t<decltype(&Op::store_tile)>> : std::true_type {};

// Tile epilogue: the finished accumulator range leaves as contiguous stores.
template <class Op, class Acc>
__device__ __forceinline__ void tile_end(const Op& op, const TileRange& t, const Acc* acc) {
    if constexpr (Planes<Op>::value == 0) return;
    __syncthreads();
    if constexpr (HasStoreTile<Op>::value) {
        op.store_tile(t.base, acc, t.size);
    } else {
        for (int i = threadIdx.x; i < t.size; i += kW) {
            if constexpr (Planes<Op>::value > 1)
                op.store_planes(t.base + i, acc + i, t.size);
            else
                op.store(t.base + i, acc[i]);
        }
    }
}

#pragma clang fp contract(fast)

}  // namespace csc
}  // namespace mpk

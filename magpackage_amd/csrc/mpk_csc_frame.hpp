// Closed-form CSC assembly (FP64 / FP32_LOCAL modes): element arithmetic in a
// node frame (lo, hi, o2, o3) whose edge (0,1) is the column's edge, and the
// two drivers that feed it.
//
//   frame driver: the column's incidences in element order; per incidence a
//                 16-byte record (element, local edge, orientation bits, row
//                 positions), the element's node ids T[e] and its two
//                 opposite nodes are gathered.
//   ring driver:  the incidences in rotation order around the column edge,
//                 so consecutive elements share a face: each record carries
//                 the element and the one new opposite node (24-bit ids), and
//                 no T[e] gather is needed.  Used when every column is a
//                 simple ring or chain (manifold meshes).
//
// Both keep the lo/hi endpoint coordinates of the column edge in registers
// for the whole column.  The frame driver runs a software pipeline over the
// incidences; the ring driver issues the gathers of kRingBatch incidences at
// once (on gfx950 the kernel is bound by L1 miss latency, so independent
// gathers in flight per wave matter more than registers: 3 measured best
// against 2, 4, 5, 6 and the pipelined form on a 12.6M-tet box).
// Summation order per CSC entry is the driver's incidence order (fixed, so
// results are deterministic); the FP64 closed forms agree with the
// reference-order arithmetic to <= 1e-12 relative.
#pragma once

#include "mpk_csc_common.hpp"

namespace mpk {
namespace csc {

// Closed-form arithmetic contracts a * b + c to an FMA only within one
// expression (not across statements, as -ffp-contract=fast allows): the
// rounding then follows the source alone, so two kernels that inline the
// same functions (e.g. the K-only and the K + M passes) give the same bits.
#pragma clang fp contract(on)

constexpr int kRingBatch = 3;

// In-frame row positions pos'_0..pos'_5 (bytes) of the 6 frame edges.
struct FPos {
    uint32_t lo, hi;
};
__device__ __forceinline__ int fpos(const FPos& p, int rp) {
    return rp < 4 ? (int)((p.lo >> (8 * rp)) & 0xffu) : (int)((p.hi >> (8 * (rp - 4))) & 0xffu);
}
__device__ __forceinline__ double fsign(uint32_t fbits, int rp) {
    return ((fbits >> rp) & 1u) ? 1.0 : -1.0;
}

// Frame record: x = element, y = c | f << 8 | pos'0 << 16 | pos'1 << 24,
// z = pos'2..pos'5.
__device__ __forceinline__ uint32_t frec_elem(uint4 r) { return r.x; }
__device__ __forceinline__ int frec_c(uint4 r) { return (int)(r.y & 0xffu); }
__device__ __forceinline__ uint32_t frec_f(uint4 r) { return (r.y >> 8) & 0x3fu; }
__device__ __forceinline__ FPos frec_pos(uint4 r) { return FPos{(r.y >> 16) | (r.z << 16), r.z >> 16}; }

// Ring record: x = element | f << 24, y = o_new | pos'0 << 24,
// z = pos'1..pos'4, w = pos'5 | o_start << 8.
__device__ __forceinline__ uint32_t rrec_elem(uint4 r) { return r.x & 0xffffffu; }
__device__ __forceinline__ uint32_t rrec_f(uint4 r) { return (r.x >> 24) & 0x3fu; }
__device__ __forceinline__ int rrec_new(uint4 r) { return (int)(r.y & 0xffffffu); }
__device__ __forceinline__ int rrec_start(uint4 r) { return (int)(r.w >> 8); }
__device__ __forceinline__ FPos rrec_pos(uint4 r) {
    return FPos{(r.y >> 24) | (r.z << 8), (r.z >> 24) | ((r.w & 0xffu) << 8)};
}

__device__ __forceinline__ Vec3 ld_node(const double4* __restrict__ P4, int node) {
    const double2 xy = *reinterpret_cast<const double2*>(&P4[node].x);
    const double z = P4[node].z;
    return Vec3{xy.x, xy.y, z};
}

// Adds v to an LDS accumulator entry with one ds_add_f64 (no return value):
// no read back and no wait, where `a += v` costs a read, a wait and a write.
// Each entry belongs to one lane's column, so the additions to it happen in
// that lane's program order, the order (and the rounding) of `a += v`.
// Column tile, ring driver, 4M-tet Newton set: 0.718-0.725 -> 0.712-0.715 ms
// FP32_LOCAL; K + j w M on the 128^3 box: 1.364 -> 1.330 ms.
__device__ __forceinline__ void lds_add(double* a, double v) {
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct FrameNodes {
    Vec3 x[4];
};

// Geometry of one incidence in the frame.
struct FrameGeom {
    Vec3 u[6];   // opposite-edge vectors: J c_r' = 2 u_r'
    Vec3 X[4];   // det * grad lambda_i (rows of adj(J))
    double det;  // signed determinant in the frame
};

__device__ __forceinline__ FrameGeom frame_geom(const FrameNodes& nf, bool need_x) {
    FrameGeom g;
    const Vec3 x0 = nf.x[0], x1 = nf.x[1], x2 = nf.x[2], x3 = nf.x[3];
    const Vec3 d1 = sub(x1, x0), d2 = sub(x2, x0), d3 = sub(x3, x0);
    g.u[0] = sub(x3, x2);
    g.u[1] = sub(x1, x3);
    g.u[2] = sub(x2, x1);
    g.u[3] = d3;
    g.u[4] = Vec3{-d2.x, -d2.y, -d2.z};
    g.u[5] = d1;
    const Vec3 X1 = cross(d2, d3);
    g.det = dot(d1, X1);
    if (need_x) {
        const Vec3 X2 = cross(d3, d1), X3 = cross(d1, d2);
        g.X[1] = X1;
        g.X[2] = X2;
        g.X[3] = X3;
        g.X[0] = Vec3{-(X1.x + X2.x + X3.x), -(X1.y + X2.y + X3.y), -(X1.z + X2.z + X3.z)};
    }
    return g;
}

// Mass column 0 in the frame (without sigma/|det| and signs): the rule's sum
// of W_r' . W_0 = Q_ak G_bl - Q_al G_bk - Q_bk G_al + Q_bl G_ak for
// r' = (a,b), (k,l) = (0,1), with symmetric moments qd / qo.
template <class R, class V>
__device__ __forceinline__ void mass_col_t(const V (&X)[4], R qd, R qo, R (&m)[6]) {
    auto d = [](const V& a, const V& b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    const R H00 = d(X[0], X[0]), H01 = d(X[0], X[1]), H11 = d(X[1], X[1]);
    const R H20 = d(X[2], X[0]), H21 = d(X[2], X[1]), H30 = d(X[3], X[0]), H31 = d(X[3], X[1]);
    m[0] = qd * (H11 + H00) - (qo + qo) * H01;
    m[1] = qd * H21 - qo * H20 - qo * H01 + qo * H00;
    m[2] = qd * H31 - qo * H30 - qo * H01 + qo * H00;
    m[3] = qo * H21 - qd * H20 - qo * H11 + qo * H01;
    m[4] = qo * H31 - qd * H30 - qo * H11 + qo * H01;
    m[5] = qo * (H31 - H30 - H21 + H20);
}

template <int kPrec>
__device__ __forceinline__ void mass_col(const FrameGeom& g, const CscQuad& q, double (&m)[6]) {
    if (kPrec == MPK_PREC_FP32_LOCAL) {
        const Vec3f X[4] = {tof(g.X[0]), tof(g.X[1]), tof(g.X[2]), tof(g.X[3])};
        float mf[6];
        mass_col_t<float, Vec3f>(X, (float)q.qd, (float)q.qo, mf);
#pragma unroll
        for (int r = 0; r < 6; ++r) m[r] = (double)mf[r];
    } else {
        mass_col_t<double, Vec3>(g.X, q.qd, q.qo, m);
    }
}

// Column 0 of the curl-curl Gram (u_r' . u_0) in FP64 or FP32.
template <int kPrec>
__device__ __forceinline__ double curl_dot(const FrameGeom& g, int rp) {
    if (kPrec == MPK_PREC_FP32_LOCAL) return (double)dotf(tof(g.u[rp]), tof(g.u[0]));
    return dot(g.u[rp], g.u[0]);
}

struct NoIds {};

// ------------------------------------------------------------------ ops --
// Each op: Acc (LDS accumulator type), Elem (per-element data gathered one
// stage ahead), Ids (per-incidence gathers two stages ahead), Col (per-column
// state), and compute() on the frame nodes.

template <int kPrec>
struct FrameCurl {
    using Acc = double;
    using Ids = NoIds;
    using Elem = double;  // nu
    struct Col {};
    const double* __restrict__ nu;
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { st_out(out + p, a); }
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        store_run(out, base, size, [&](int i) { return acc[i]; });
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return nu[e]; }
    __device__ void compute(const FrameNodes& nodes, const Elem& nuv, uint32_t f, const FPos& pos,
                            Acc* acc, Col&) const {
        const FrameGeom g = frame_geom(nodes, false);
        const double scale = 4.0 * qa.wsum * nuv / fabs(g.det) * fsign(f, 0);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = fpos(pos, rp);
            const double d = curl_dot<kPrec>(g, rp);
            if (p != 0xff) lds_add(&acc[p], d * (scale * fsign(f, rp)));
        }
    }
};

template <int kPrec>
struct FrameMass {
    using Acc = double;
    using Ids = NoIds;
    using Elem = double;  // sigma
    struct Col {};
    const double* __restrict__ coeffs;  // NULL: sigma = 1
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { st_out(out + p, a); }
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        store_run(out, base, size, [&](int i) { return acc[i]; });
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return coeffs ? coeffs[e] : 1.0; }
    __device__ void compute(const FrameNodes& nodes, const Elem& sigma, uint32_t f,
                            const FPos& pos, Acc* acc, Col&) const {
        if (!(fabs(sigma) > 1e-12)) return;  // zero block (mass_kernel_mex.cpp:69)
        const FrameGeom g = frame_geom(nodes, true);
        const double scale = sigma / fabs(g.det) * fsign(f, 0);
        double m[6];
        mass_col<kPrec>(g, qa, m);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = fpos(pos, rp);
            if (p != 0xff) lds_add(&acc[p], m[rp] * (scale * fsign(f, rp)));
        }
    }
};

template <int kPrec>
struct FrameKM {
    using Acc = double2;
    static constexpr int kRing = 6;  // incidences per gather round (ring driver)
    using Ids = NoIds;
    using Elem = double2;  // (nu, sigma)
    struct Col {};
    const double* __restrict__ nu;
    const double* __restrict__ sigma;
    double2* __restrict__ out;
    double omega;
    CscQuad qk, qm;
    int acc_bytes_max;
    __device__ static Acc zero() { return make_double2(0.0, 0.0); }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const {
        // K + 1j*omega*M: real part K, imaginary part omega * (sum of M)
        st_out(out + p, make_double2(a.x, omega * a.y));
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return make_double2(nu[e], sigma[e]); }
    __device__ void compute(const FrameNodes& nodes, const Elem& el, uint32_t f, const FPos& pos,
                            Acc* acc, Col&) const {
        const bool has_m = fabs(el.y) > 1e-12;
        const FrameGeom g = frame_geom(nodes, has_m);
        const double inv = 1.0 / fabs(g.det);
        const double f0 = fsign(f, 0);
        const double ks = 4.0 * qk.wsum * el.x * inv * f0;
        const double ms = el.y * inv * f0;
        double m[6] = {0, 0, 0, 0, 0, 0};
        if (has_m) mass_col<kPrec>(g, qm, m);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = fpos(pos, rp);
            const double d = curl_dot<kPrec>(g, rp);
            const double s = fsign(f, rp);
            if (p != 0xff) {
                lds_add(&acc[p].x, d * (ks * s));
                if (has_m) lds_add(&acc[p].y, m[rp] * (ms * s));
            }
        }
    }
};

// Newton quantities read the element state written by k_elem_state once per
// call: B (the element's constant flux density, from A), nu(|B|^2), dnu and
// sigma.  With B known, the frame curls C'_r' = 2 u_r' / det give
// B.C'_r' = (2/det) B.u_r' directly; no DOF or solution gathers per visit.
struct ElemState {
    double4 s0;  // Bx, By, Bz, nu
    double2 s1;  // dnu, sigma
};
// A state view gathers an element's state as it is stored (Elem / Bnu, kept
// in that form while the gathers of a ring batch are in flight) and widens it
// to the FP64 arithmetic's types at the point of use (full / bnu).
struct StateView {
    using Elem = ElemState;
    using Bnu = double4;
    using S0 = double4;
    using S1 = double2;
    const double4* __restrict__ s0;
    const double2* __restrict__ s1;
    __device__ Elem load(uint32_t e) const { return Elem{s0[e], s1[e]}; }
    __device__ Bnu load_bnu(uint32_t e) const { return s0[e]; }
    __device__ static ElemState full(const Elem& r) { return r; }
    __device__ static double4 bnu(const Bnu& b) { return b; }
};
// FP32-local element state: the same fields rounded to FP32, 24 B per element
// (half the gather and HBM bytes of the FP64 state), held as FP32 in
// registers (half the registers per gather in flight) and widened to FP64
// where the arithmetic reads it, which is then the FP64-state code.
struct StateView32 {
    struct Raw {
        float4 a;
        float2 b;
    };
    using Elem = Raw;
    using Bnu = float4;
    using S0 = float4;
    using S1 = float2;
    const float4* __restrict__ s0;
    const float2* __restrict__ s1;
    __device__ Elem load(uint32_t e) const { return Raw{s0[e], s1[e]}; }
    __device__ Bnu load_bnu(uint32_t e) const { return s0[e]; }
    __device__ static ElemState full(const Elem& r) {
        return ElemState{make_double4(r.a.x, r.a.y, r.a.z, r.a.w), make_double2(r.b.x, r.b.y)};
    }
    __device__ static double4 bnu(const Bnu& a) { return make_double4(a.x, a.y, a.z, a.w); }
};

// Newton Jacobian J = K(nu) + 2 dnu (B.C_r)(B.C_c) and residual R = K(nu) a
// (assemble_jacobian_kernel_mex.cpp:94-131 per visit, in the frame).
template <int kPrec, bool kCalcJ, class SV = StateView>
struct FrameJac {
    using Acc = double;
    using Ids = NoIds;
    using Elem = typename SV::Elem;
    struct Col {
        double r;
    };
    static constexpr int kPlanes = kCalcJ ? 1 : 0;  // residual only: no LDS tile
    SV st;
    double* __restrict__ out;
    double* __restrict__ res;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{0.0}; }
    __device__ void col_end(int64_t j, Col& cs) const { res[j] = cs.r; }
    __device__ void store(int64_t p, Acc a) const {
        if (kCalcJ) st_out(out + p, a);
    }
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        if (kCalcJ) store_run(out, base, size, [&](int i) { return acc[i]; });
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return st.load(e); }
    __device__ void compute(const FrameNodes& nodes, const Elem& raw, uint32_t f,
                            const FPos& pos, Acc* acc, Col& cs) const {
        const ElemState el = SV::full(raw);
        const FrameGeom g = frame_geom(nodes, false);
        const Vec3 Bv{el.s0.x, el.s0.y, el.s0.z};
        const double nu = el.s0.w, dnu = el.s1.x;
        const double two_inv = 2.0 / g.det;
        // w|det| summed over the rule; C'.C' = 4 u.u / det^2
        const double wd = qa.wsum * fabs(g.det);
        const double f0 = fsign(f, 0);
        const double BC0 = two_inv * dot(Bv, g.u[0]);
        cs.r += nu * BC0 * wd * f0;
        if (kCalcJ) {
            const double cc = 4.0 / (g.det * g.det);
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = fpos(pos, rp);
                const double BCr = two_inv * dot(Bv, g.u[rp]);
                const double uu = curl_dot<kPrec>(g, rp);
                const double val = nu * (cc * uu) + 2.0 * dnu * BCr * BC0;
                if (p != 0xff) lds_add(&acc[p], val * wd * (fsign(f, rp) * f0));
            }
        }
    }
};

// Fused Newton set: in one pass over the plan, the secant stiffness K(nu(B))
// (the K assembleStiffness builds from the per-element nu the B-H curve gives
// at B), the sigma-mass M, the Jacobian J and the residual R.  Outputs may be
// NULL (skipped).  LDS planes K, D = J - K (the 2 dnu (B.C)(B.C) term) and,
// when kWithM, M: visits of linear elements (dnu = 0) touch only K and visits
// of non-conducting ones (sigma = 0) skip M; J = K + D at the store.  Without
// M the tile needs 16 B of LDS per entry instead of 24 (more waves per CU).
template <int kPrec, bool kWithM, class SV = StateView>
struct FrameNewton {
    using Acc = double;
    static constexpr int kRing = 6;  // incidences per gather round (ring driver)
    using Ids = NoIds;
    using Elem = typename SV::Elem;
    struct Col {
        double r;
    };
    static constexpr int kPlanes = kWithM ? 3 : 2;
    static constexpr int kD = kWithM ? 2 : 1;  // plane of D
    SV st;
    double* __restrict__ out_k;
    double* __restrict__ out_m;
    double* __restrict__ out_j;
    double* __restrict__ res;
    CscQuad qk, qm;
    int acc_bytes_max;
    int stride;  // plane stride (the block's tile size), set by tile_begin
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{0.0}; }
    __device__ void col_end(int64_t j, Col& cs) const {
        if (res) res[j] = cs.r;
    }
    __device__ void store_planes(int64_t p, const Acc* a, int size) const {
        const double k = a[0];
        if (out_k) st_out(out_k + p, k);
        if (kWithM) st_out(out_m + p, a[size]);
        if (out_j) st_out(out_j + p, k + a[kD * size]);
    }
    // array-outer: each output's tile range is written as one contiguous run
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        if (out_k) store_run(out_k, base, size, [&](int i) { return acc[i]; });
        if (kWithM) store_run(out_m, base, size, [&](int i) { return acc[size + i]; });
        if (out_j)
            store_run(out_j, base, size, [&](int i) { return acc[i] + acc[kD * size + i]; });
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return st.load(e); }
    __device__ void compute(const FrameNodes& nodes, const Elem& raw, uint32_t f,
                            const FPos& pos, Acc* acc, Col& cs) const {
        const ElemState el = SV::full(raw);
        const double sigma = el.s1.y;
        const bool has_m = kWithM && fabs(sigma) > 1e-12;
        const FrameGeom g = frame_geom(nodes, has_m);
        const Vec3 Bv{el.s0.x, el.s0.y, el.s0.z};
        const double nu = el.s0.w, dnu = el.s1.x;
        const double two_inv = 2.0 / g.det;
        const double inv = 1.0 / fabs(g.det);
        const double wd = qk.wsum * fabs(g.det);
        const double f0 = fsign(f, 0);
        const double BC0 = two_inv * dot(Bv, g.u[0]);
        cs.r += nu * BC0 * wd * f0;
        const double ks = 4.0 * qk.wsum * nu * inv * f0;  // = nu * (4/det^2) * wd
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = fpos(pos, rp);
            if (p != 0xff) lds_add(&acc[p], curl_dot<kPrec>(g, rp) * (ks * fsign(f, rp)));
        }
        if (has_m) {
            const double ms = sigma * inv * f0;
            double m[6];
            mass_col<kPrec>(g, qm, m);
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = fpos(pos, rp);
                if (p != 0xff) lds_add(&acc[stride + p], m[rp] * (ms * fsign(f, rp)));
            }
        }
        if (out_j && dnu != 0.0) {
            const double js = 2.0 * dnu * BC0 * wd * f0;
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = fpos(pos, rp);
                if (p != 0xff)
                    lds_add(&acc[kD * stride + p], (two_inv * dot(Bv, g.u[rp])) * (js * fsign(f, rp)));
            }
        }
    }
};

// The fused Newton set on a tile none of whose elements is conducting
// (|sigma| <= 1e-12) or nonlinear at its B (dnu = 0): no M or D term reaches
// such a tile, so it needs the K plane only (8 B of LDS per entry: about
// three times the waves per CU of FrameNewton) and gathers (B, nu) alone.  The
// store writes M = 0 and J = K + 0, the values FrameNewton stores for the
// same tile.
template <int kPrec, class SV>
struct FrameNewtonLin {
    using Acc = double;
    static constexpr int kRing = 6;  // incidences per gather round (ring driver)
    using Ids = NoIds;
    using Elem = typename SV::Bnu;  // Bx, By, Bz, nu as stored
    struct Col {
        double r;
    };
    SV st;
    double* __restrict__ out_k;
    double* __restrict__ out_m;
    double* __restrict__ out_j;
    double* __restrict__ res;
    CscQuad qk;
    int acc_bytes_max;
    // Optional: the list of class-1 tiles, built here instead of by a kernel
    // of its own.  Block b lists the tiles among [64 b, 64 b + 64) whose flag
    // holds list_epoch (ascending), the blocks' ranges appended in arrival
    // order, as k_tile_list does; list_flag holds the finished classes of
    // every tile (written by an earlier launch) and *list_count starts at 0.
    const int32_t* __restrict__ list_flag = nullptr;
    int32_t* __restrict__ list_out = nullptr;
    int32_t* __restrict__ list_count = nullptr;
    int64_t list_ntiles = 0;
    int32_t list_epoch = 1;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {
        if (!list_flag) return;
        const int64_t t = (int64_t)blockIdx.x * kW + threadIdx.x;
        if ((int64_t)blockIdx.x * kW >= list_ntiles) return;  // uniform per block
        const bool on = t < list_ntiles && list_flag[t] == list_epoch;
        const unsigned long long m = __ballot(on);
        if (m == 0ull) return;
        const int lane = threadIdx.x;
        int base = 0;
        if (lane == 0) base = atomicAdd(list_count, __popcll(m));
        base = __shfl(base, 0);
        if (on) list_out[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)t;
    }
    __device__ Col col_begin() const { return Col{0.0}; }
    __device__ void col_end(int64_t j, Col& cs) const {
        if (res) res[j] = cs.r;
    }
    __device__ void store(int64_t p, Acc k) const {
        if (out_k) st_out(out_k + p, k);
        if (out_m) st_out(out_m + p, 0.0);
        if (out_j) st_out(out_j + p, k + 0.0);
    }
    // array-outer: each output's tile range is written as one contiguous run
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        if (out_k) store_run(out_k, base, size, [&](int i) { return acc[i]; });
        if (out_m) store_run(out_m, base, size, [&](int) { return 0.0; });
        if (out_j) store_run(out_j, base, size, [&](int i) { return acc[i] + 0.0; });
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids&) const { return st.load_bnu(e); }
    __device__ void compute(const FrameNodes& nodes, const Elem& raw, uint32_t f,
                            const FPos& pos, Acc* acc, Col& cs) const {
        const double4 el = SV::bnu(raw);
        const FrameGeom g = frame_geom(nodes, false);
        const Vec3 Bv{el.x, el.y, el.z};
        const double nu = el.w;
        const double two_inv = 2.0 / g.det;
        const double inv = 1.0 / fabs(g.det);
        const double wd = qk.wsum * fabs(g.det);
        const double f0 = fsign(f, 0);
        const double BC0 = two_inv * dot(Bv, g.u[0]);
        cs.r += nu * BC0 * wd * f0;
        const double ks = 4.0 * qk.wsum * nu * inv * f0;
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = fpos(pos, rp);
            if (p != 0xff) lds_add(&acc[p], curl_dot<kPrec>(g, rp) * (ks * fsign(f, rp)));
        }
    }
};

// -------------------------------------------------------------- drivers --
// Frame driver, element order.  Per lane, a 4-stage pipeline:
//   stage 3: frame record                            for k + 3
//   stage 2: T[e] and the op's Ids                   for k + 2
//   stage 1: the 2 opposite nodes and Elem           for k + 1
//   stage 0: arithmetic and LDS accumulation         for k
template <class Op>
__global__ void __launch_bounds__(kW) k_csc_frame(TileCtx tc, Op op) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Acc = typename Op::Acc;
    using Ids = typename Op::Ids;
    using Elem = typename Op::Elem;
    Acc* acc = reinterpret_cast<Acc*>(smem);
    const TileRange tr = tile_begin(tc, op, acc, smem);
    const int lane = threadIdx.x;
    const int64_t j = tr.j0 + lane;
    if (j < tr.jend) {
        const int cbase = tc.colptr[j];
        Acc* col = acc + (cbase - tr.base);
        typename Op::Col cs = op.col_begin();
        const int cnt = tc.inc_ptr[j + 1] - tc.inc_ptr[j];
        if (cnt > 0) {
            // k-major records: incidence k of this column at rec[k * kW]
            const uint4* __restrict__ rec = tc.inc_frame + tc.frame_base[tr.tile_id] + lane;
            const uint4 z4 = make_uint4(0, 0, 0, 0);
            uint4 r0 = rec[0], r1 = z4, r2 = z4;
            if (1 < cnt) r1 = rec[kW];
            if (2 < cnt) r2 = rec[2 * kW];
            int4 t0 = tc.T[frec_elem(r0)], t1 = t0;
            Ids d0 = op.ids(tc, cbase, frec_pos(r0)), d1 = d0;
            if (1 < cnt) {
                t1 = tc.T[frec_elem(r1)];
                d1 = op.ids(tc, cbase, frec_pos(r1));
            }
            // the column edge's endpoints, common to all its incidences
            const int c0 = frec_c(r0);
            const int na = sel6(c0, t0.x, t0.x, t0.x, t0.y, t0.y, t0.z);
            const int nb = sel6(c0, t0.y, t0.z, t0.w, t0.z, t0.w, t0.w);
            const Vec3 xlo = ld_node(tc.P4, na <= nb ? na : nb);
            const Vec3 xhi = ld_node(tc.P4, na <= nb ? nb : na);
            auto opp2 = [&](int4 t, int c) {
                return ld_node(tc.P4, sel6(c, t.z, t.y, t.y, t.x, t.x, t.x));
            };
            auto opp3 = [&](int4 t, int c) {
                return ld_node(tc.P4, sel6(c, t.w, t.w, t.z, t.w, t.z, t.y));
            };
            Vec3 a2 = opp2(t0, c0), a3 = opp3(t0, c0);
            Elem e0 = op.elem(frec_elem(r0), d0);
            for (int k = 0; k < cnt; ++k) {
                Vec3 b2 = a2, b3 = a3;
                Elem e1 = e0;
                if (k + 1 < cnt) {
                    b2 = opp2(t1, frec_c(r1));
                    b3 = opp3(t1, frec_c(r1));
                    e1 = op.elem(frec_elem(r1), d1);
                }
                int4 t2 = t1;
                Ids d2 = d1;
                if (k + 2 < cnt) {
                    t2 = tc.T[frec_elem(r2)];
                    d2 = op.ids(tc, cbase, frec_pos(r2));
                }
                uint4 r3 = z4;
                if (k + 3 < cnt) r3 = rec[(k + 3) * kW];
                op.compute(FrameNodes{{xlo, xhi, a2, a3}}, e0, frec_f(r0), frec_pos(r0), col, cs);
                a2 = b2; a3 = b3; e0 = e1;
                r0 = r1; r1 = r2; r2 = r3;
                t1 = t2; d1 = d2;
            }
        }
        op.col_end(j, cs);
    }
    tile_end(op, tr, acc);
}

// Ring driver, batched.  The incidences of a column are taken kB at a time:
// the batch's records, then its Ids, then its new nodes and Elem are all
// issued before any arithmetic, so a wave keeps up to kB independent gathers
// in flight per level instead of one.  Load indices are clamped to the last
// incidence (always valid addresses, no branches in the load stream); the
// arithmetic of slots past the column's count is skipped.
// Every index of a tile depends on the tile and the lane only: the loads go
// out together, one memory latency before the records instead of a chain of
// dependent ones.
// kC selects the record format: compact (8-byte records + a 16-byte column
// head; positions of the shared face and orientation bits carried from one
// incidence to the next) or the 16-byte ring records.
// Streaming loads (non-temporal hint) of the compact ring records and column
// heads: each is read once per pass, and the hint keeps them from displacing
// the gathered node and element data in L2.  Same bits as a plain load.
// 128^3-box K + jwM: 1.313 -> 1.308 ms (FP64), 1.248 -> 1.238 ms (FP32-local);
// 4M-tet Newton set 0.665-0.671 -> 0.656-0.657 ms, on one box.
__device__ __forceinline__ uint2 ld_stream(const uint2* p) {
    const unsigned long long v =
        __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
    return make_uint2((unsigned)v, (unsigned)(v >> 32));
}
__device__ __forceinline__ uint4 ld_stream(const uint4* p) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

struct RingIdx {
    int cbase, cnt;  // column's first value index, incidences
    int lo, hi;      // column edge's endpoints
    uint4 head;      // compact format: the column head
    int base, top, fbase;
    int32_t cls;
};
template <bool kC>
__device__ __forceinline__ RingIdx ring_load(const TileCtx& tc, const int2* __restrict__ ends,
                                             int64_t tile_id, const int32_t* __restrict__ sel,
                                             int32_t want, int32_t epoch) {
    const int64_t j0 = tile_id * kW;
    const int64_t jend = j0 + kW < tc.n ? j0 + kW : tc.n;
    const int64_t j = j0 + threadIdx.x;
    const int64_t jc = j < jend ? j : jend - 1;  // a valid column for every lane
    RingIdx ix;
    ix.base = tc.colptr[j0];
    ix.top = tc.colptr[jend];
    ix.fbase = tc.frame_base[tile_id];
    // class 1: the tile's flag holds this call's epoch
    ix.cls = sel ? (sel[tile_id] == epoch ? 1 : 0) : want;
    if constexpr (kC) {
        const uint4 h = ld_stream(tc.col_head + jc);
        ix.head = h;
        ix.cbase = ix.base + (int)(h.w & 0xffffu);
        ix.cnt = (int)((h.w >> 16) & 0xffu);
        ix.lo = (int)(h.x & 0xffffffu);
        ix.hi = (int)(h.y & 0xffffffu);
    } else {
        ix.cbase = tc.colptr[jc];
        ix.cnt = tc.inc_ptr[jc + 1] - tc.inc_ptr[jc];
        const int2 lh = ends[jc];
        ix.lo = lh.x;
        ix.hi = lh.y;
        ix.head = make_uint4(0, 0, 0, 0);
    }
    return ix;
}

__device__ __forceinline__ uint4 ldrec(const uint4* p) { return *p; }
__device__ __forceinline__ uint2 ldrec(const uint2* p) { return ld_stream(p); }

// Compact record fields (layout in mpk_plan.hpp).
__device__ __forceinline__ uint32_t r8_elem(uint2 r) { return r.x & 0xffffffu; }
__device__ __forceinline__ int r8_new(uint2 r, int lo) {
    return lo + ((int)(r.y << 10) >> 10);  // 22-bit two's complement offset from lo
}
__device__ __forceinline__ uint32_t r8_pos6(uint32_t v) { return v == 63u ? 0xffu : v; }

// One tile of the ring driver from its loaded indices.  A tile of another
// class is not branched around (the compiler would hoist that branch above
// the index loads): it runs with no live lanes and an empty range, so it
// writes nothing.
template <class Op, int kB, bool kC>
__device__ __forceinline__ void ring_run(const TileCtx& tc, Op& op, unsigned char* smem,
                                         int64_t tile_id, const RingIdx& ix, int32_t want) {
    using Acc = typename Op::Acc;
    using Ids = typename Op::Ids;
    using Elem = typename Op::Elem;
    using Rec = std::conditional_t<kC, uint2, uint4>;
    static_assert(!kC || std::is_same<Ids, NoIds>::value,
                  "compact ring records carry no per-incidence positions for Ids");
    Acc* acc = reinterpret_cast<Acc*>(smem);
    const int lane = threadIdx.x;
    TileRange tr;
    tr.tile_id = tile_id;
    tr.j0 = tile_id * kW;
    tr.jend = tr.j0 + kW < tc.n ? tr.j0 + kW : tc.n;
    const int64_t j = tr.j0 + lane;
    const int cbase = ix.cbase;
    const int base = ix.base, top = ix.top, fbase = ix.fbase;
    const bool mine = ix.cls == want;
    const bool live = mine && j < tr.jend;
    tr.base = base;
    tr.size = mine ? top - base : 0;
    // The first gather round's records go out before the LDS accumulator is
    // cleared, for every lane: load indices are clamped to the column's last
    // record, and to record row 0 when the lane has none (every column of a
    // plan has at least one incidence, so each tile's record block has at
    // least one row of kW records; a lane without records never uses them).
    const int cnt = live ? ix.cnt : 0;
    const int klast = cnt > 0 ? cnt - 1 : 0;
    const Rec* __restrict__ rec;
    if constexpr (kC)
        rec = tc.inc_ring8 + fbase + lane;
    else
        rec = tc.inc_ring + fbase + lane;
    Rec r[kB];
#pragma unroll
    for (int i = 0; i < kB; ++i) r[i] = ldrec(rec + (i < klast ? i : klast) * kW);
    tile_zero(op, acc, smem, tr);
    if (live) {
        Acc* col = acc + (cbase - base);
        typename Op::Col cs = op.col_begin();
        if (cnt > 0) {
            const int lo = ix.lo, hi = ix.hi;
            const Vec3 xlo = ld_node(tc.P4, lo), xhi = ld_node(tc.P4, hi);
            // compact format: node id of o_prev and the positions carried
            // along the ring (pos'_0 fixed, pos'_1 / pos'_3 from the last
            // record's pos'_2 / pos'_4)
            int oprev = 0;
            uint32_t p0 = 0, p1 = 0, p3 = 0;
            if constexpr (kC) {
                oprev = (int)(ix.head.z & 0xffffffu);
                p0 = ix.head.x >> 24;
                p1 = ix.head.y >> 24;
                p3 = ix.head.z >> 24;
            } else {
                oprev = rrec_start(r[0]);
            }
            Vec3 xprev = ld_node(tc.P4, oprev);
            for (int k0 = 0; k0 < cnt; k0 += kB) {
                if (k0 > 0) {
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        const int k = k0 + i < cnt ? k0 + i : cnt - 1;
                        r[i] = ldrec(rec + k * kW);
                    }
                }
                Ids d[kB];
                int onew[kB];
                uint32_t elem[kB];
#pragma unroll
                for (int i = 0; i < kB; ++i) {
                    if constexpr (kC) {
                        d[i] = Ids{};
                        onew[i] = r8_new(r[i], lo);
                        elem[i] = r8_elem(r[i]);
                    } else {
                        d[i] = op.ids(tc, cbase, rrec_pos(r[i]));
                        onew[i] = rrec_new(r[i]);
                        elem[i] = rrec_elem(r[i]);
                    }
                }
                Vec3 xn[kB];
                Elem el[kB];
#pragma unroll
                for (int i = 0; i < kB; ++i) {
                    xn[i] = ld_node(tc.P4, onew[i]);
                    el[i] = op.elem(elem[i], d[i]);
                }
#pragma unroll
                for (int i = 0; i < kB; ++i) {
                    if (k0 + i < cnt) {
                        if constexpr (kC) {
                            const uint2 q = r[i];
                            const uint32_t p2 = r8_pos6((q.x >> 24) & 63u);
                            const uint32_t p4 = r8_pos6((q.y >> 22) & 63u);
                            const uint32_t p5 = r8_pos6((q.x >> 30) | ((q.y >> 28) << 2));
                            const int on = onew[i];
                            const uint32_t f = 1u | (uint32_t)(lo < oprev) << 1 |
                                               (uint32_t)(lo < on) << 2 |
                                               (uint32_t)(hi < oprev) << 3 |
                                               (uint32_t)(hi < on) << 4 |
                                               (uint32_t)(oprev < on) << 5;
                            op.compute(FrameNodes{{xlo, xhi, xprev, xn[i]}}, el[i], f,
                                       FPos{p0 | p1 << 8 | p2 << 16 | p3 << 24, p4 | p5 << 8},
                                       col, cs);
                            p1 = p2;
                            p3 = p4;
                            oprev = on;
                        } else {
                            // bit 0 (lo -> hi) is 1 in every ring record; set
                            // as a constant, both formats compile the same
                            // arithmetic (and round identically)
                            op.compute(FrameNodes{{xlo, xhi, xprev, xn[i]}}, el[i],
                                       rrec_f(r[i]) | 1u, rrec_pos(r[i]), col, cs);
                        }
                        xprev = xn[i];
                    }
                }
            }
        }
        op.col_end(j, cs);
    }
    tile_end(op, tr, acc);
}

template <class Op, int kB, bool kC>
__device__ __forceinline__ void ring_tile(const TileCtx& tc, const int2* __restrict__ ends, Op& op,
                                          unsigned char* smem, int64_t tile_id,
                                          const int32_t* __restrict__ sel = nullptr,
                                          int32_t want = 0, int32_t epoch = 1) {
    const RingIdx ix = ring_load<kC>(tc, ends, tile_id, sel, want, epoch);
    ring_run<Op, kB, kC>(tc, op, smem, tile_id, ix, want);
}

// sel (optional): per-tile class flags; a block whose tile's flag differs
// from `want` returns at once (the tiles of a plan split over two launches).
template <class Op, int kB, bool kC>
__global__ void __launch_bounds__(kW) k_csc_ring_batch(TileCtx tc, const int2* __restrict__ ends,
                                                       Op op, const int32_t* __restrict__ sel,
                                                       int32_t want, int32_t epoch) {
    extern __shared__ __align__(16) unsigned char smem[];
    ring_tile<Op, kB, kC>(tc, ends, op, smem, xcd_tile(blockIdx.x, gridDim.x), sel, want,
                          epoch);
}

// Tiles from a device list (persistent grid): block b runs on XCD b % 8; the
// blocks of XCD x walk the contiguous list range [x c / 8, (x+1) c / 8),
// c = *count, so neighbouring tiles of the list share that XCD's L2.
template <class Op, int kB, bool kC>
__global__ void __launch_bounds__(kW) k_csc_ring_list(TileCtx tc, const int2* __restrict__ ends,
                                                      Op op, const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ count) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int64_t kXcd = 8;
    const int64_t c = *count;
    const int64_t x = blockIdx.x % kXcd;
    const int64_t nbx = ((int64_t)gridDim.x - x + kXcd - 1) / kXcd;  // blocks on XCD x
    const int64_t lo = c * x / kXcd, hi = c * (x + 1) / kXcd;
    for (int64_t p = lo + blockIdx.x / kXcd; p < hi; p += nbx) {
        ring_tile<Op, kB, kC>(tc, ends, op, smem, list[p]);
        __syncthreads();
    }
}

// Two ops over the tiles of a device list, one after the other on each tile
// (the second op's records and gathers come back from L2): the Newton set's
// listed tiles, K, J and R from the two-plane op and M from the mass op, in
// one persistent launch.  The LDS reservation is the larger of the two ops'.
template <class Op1, class Op2, int kB1, int kB2, bool kC>
__global__ void __launch_bounds__(kW) k_csc_ring_list2(TileCtx tc, const int2* __restrict__ ends,
                                                       Op1 op1, Op2 op2,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ count) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int64_t kXcd = 8;
    const int64_t c = *count;
    const int64_t x = blockIdx.x % kXcd;
    const int64_t nbx = ((int64_t)gridDim.x - x + kXcd - 1) / kXcd;  // blocks on XCD x
    const int64_t lo = c * x / kXcd, hi = c * (x + 1) / kXcd;
    for (int64_t p = lo + blockIdx.x / kXcd; p < hi; p += nbx) {
        const int64_t t = list[p];
        ring_tile<Op1, kB1, kC>(tc, ends, op1, smem, t);
        __syncthreads();
        ring_tile<Op2, kB2, kC>(tc, ends, op2, smem, t);
        __syncthreads();
    }
}

// K + j w_f M_sigma for several frequencies from one pass: the LDS tile holds
// (K, M) as FrameKM's does; the store writes one complex128 value set per
// frequency, laid out [f][nnz] (HBFEM-style harmonic blocks, frequency sweeps).
constexpr int kMaxFreq = 64;
template <int kPrec>
struct FrameKMF {
    using Acc = double2;
    using Ids = NoIds;
    using Elem = double2;
    struct Col {};
    FrameKM<kPrec> km;
    const double* __restrict__ omegas;  // [nfreq]
    int nfreq;
    int64_t nnz;
    int acc_bytes_max;
    __device__ static Acc zero() { return make_double2(0.0, 0.0); }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const {
        for (int f = 0; f < nfreq; ++f) km.out[f * nnz + p] = make_double2(a.x, omegas[f] * a.y);
    }
    // frequency-outer: each value set's tile range is written as one
    // contiguous run before the next set's.  Plain stores: this write-bound
    // pass measured 24-25 ms with them and 29.4 ms with the non-temporal hint.
    __device__ void store_tile(int64_t base, const Acc* acc, int size) const {
        for (int f = 0; f < nfreq; ++f) {
            const double w = omegas[f];
            double2* __restrict__ o = km.out + f * nnz + base;
            for (int i = threadIdx.x; i < size; i += kW) {
                const Acc a = acc[i];
                o[i] = make_double2(a.x, w * a.y);
            }
        }
    }
    __device__ Ids ids(const TileCtx&, int, const FPos&) const { return Ids{}; }
    __device__ Elem elem(uint32_t e, const Ids& d) const { return km.elem(e, typename FrameKM<kPrec>::Ids{}); }
    __device__ void compute(const FrameNodes& nodes, const Elem& el, uint32_t f, const FPos& pos,
                            Acc* acc, Col&) const {
        typename FrameKM<kPrec>::Col c{};
        km.compute(nodes, el, f, pos, acc, c);
    }
};

#pragma clang fp contract(fast)

}  // namespace csc
}  // namespace mpk

// Column-owner CSC assembly kernels (the hot path) on gfx950.
//
// Design: one wave64 block owns 64 consecutive CSC columns (edge DOFs), one
// lane per column.  The block's contiguous value range
// [colptr[j0], colptr[j0+64]) is accumulated in LDS; each lane walks the
// incidence list of its column -- the (element, local edge) pairs in
// ascending element order, built once by mpk_plan_create -- recomputes the
// element's column of the local block in registers and adds the 6 values at
// their in-column positions.  A lane only ever touches its own column, so
// there are no atomics and the summation order per entry is the triplet
// order of the reference's COO output (element-major).  The finished tile
// leaves the CU as one contiguous, coalesced store.  HBM traffic per tet is
// the mesh inputs plus the CSC values written once; the 36 COO triplets
// (864 B/tet in the reference) are never materialised.
//
// Blocks are mapped to tiles XCD-major (xcd_tile) and, per lane, incidences
// flow through a software pipeline so the dependent loads of the next
// incidences are in flight while the current one is computed.
//
// Element arithmetic, per precision mode:
//   FP64_EXACT  reference operation order, no FMA contraction (direct driver):
//               src/mex/assemble_curl_curl_kernel_mex.cpp:57-101,
//               src/mex/assemble_mass_kernel_mex.cpp:58-108,
//               src/mex/assemble_jacobian_kernel_mex.cpp:71-131.
//   FP64        closed forms in a frame where the lane's edge is local edge
//               0 (nodes permuted), FMA allowed (frame driver).  Curl-curl:
//               K_rc = 4 nu W (u_r . u_c) / |det|, u = opposite-edge vectors
//               (J c_i = 2 u_i for the Nedelec P1 reference curls).  Mass:
//               M_rc = sigma/|det| * sum Q_xy (X . X) with X_i = det*grad l_i
//               and Q the barycentric moments of the same quadrature rule.
//   FP32_LOCAL  as FP64 but the dot products and block entries in FP32;
//               differences, adjugate, determinant and accumulation in FP64.
#include <cmath>
#include <cstring>

#include "mpk_plan.hpp"

namespace mpk {
namespace {

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
    const int32_t* __restrict__ frame_base;
    const int4* __restrict__ T;
    const double4* __restrict__ P4;
    const int32_t* __restrict__ rowind;
    int64_t n;
};

__device__ __forceinline__ int pos_of(uint2 pw, int r) {
    return r < 4 ? (int)((pw.x >> (8 * r)) & 0xffu) : (int)((pw.y >> (8 * (r - 4))) & 0xffu);
}

template <class T>
__device__ __forceinline__ T sel6(int c, T v0, T v1, T v2, T v3, T v4, T v5) {
    return c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : c == 3 ? v3 : c == 4 ? v4 : v5;
}

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

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so that each XCD sweeps one contiguous range of column tiles: the
// elements and nodes shared by neighbouring tiles then stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    constexpr int64_t kXcd = 8;
    const int64_t q = nb / kXcd, r = nb % kXcd;
    const int64_t x = b % kXcd, i = b / kXcd;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// =========================================================================
// Frame path (FP64 / FP32_LOCAL closed forms)
// =========================================================================
// Frame record fields (mpk_plan.hpp): element, local edge c, orientation bits
// of the 6 frame edges and their in-column positions.
__device__ __forceinline__ uint32_t frec_elem(uint4 r) { return r.x; }
__device__ __forceinline__ int frec_c(uint4 r) { return (int)(r.y & 0xffu); }
__device__ __forceinline__ double frec_f(uint4 r, int rp) {
    return ((r.y >> (8 + rp)) & 1u) ? 1.0 : -1.0;
}
__device__ __forceinline__ int frec_pos(uint4 r, int rp) {
    return rp < 2 ? (int)((r.y >> (16 + 8 * rp)) & 0xffu) : (int)((r.z >> (8 * (rp - 2))) & 0xffu);
}

// Node coordinates in the frame (lo, hi, o1, o2).  lo/hi are the column
// edge's endpoints, shared by every incidence of the column and loaded once;
// per incidence only the two opposite nodes o1, o2 are gathered.
struct FrameNodes {
    Vec3 x[4];
};
struct FrameOpp {
    Vec3 x2, x3;
};
__device__ __forceinline__ Vec3 ld_node(const double4* __restrict__ P4, int node) {
    const double2 xy = *reinterpret_cast<const double2*>(&P4[node].x);
    const double z = P4[node].z;
    return Vec3{xy.x, xy.y, z};
}
__device__ __forceinline__ FrameOpp fetch_opp(const double4* __restrict__ P4, int4 t, int c) {
    const int n2 = sel6(c, t.z, t.y, t.y, t.x, t.x, t.x);
    const int n3 = sel6(c, t.w, t.w, t.z, t.w, t.z, t.y);
    return FrameOpp{ld_node(P4, n2), ld_node(P4, n3)};
}
// Global ids of the endpoints (lo, hi) of local edge c of element t.
__device__ __forceinline__ int2 edge_ends(int4 t, int c) {
    const int na = sel6(c, t.x, t.x, t.x, t.y, t.y, t.z);
    const int nb = sel6(c, t.y, t.z, t.w, t.z, t.w, t.w);
    return na <= nb ? make_int2(na, nb) : make_int2(nb, na);
}
__device__ __forceinline__ FrameNodes frame_nodes(const Vec3& lo, const Vec3& hi,
                                                  const FrameOpp& o) {
    return FrameNodes{{lo, hi, o.x2, o.x3}};
}

// Geometry of one incidence in the frame (lane edge == edge (0,1)).
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
__device__ __forceinline__ void mass_col(const FrameGeom& g, double qd, double qo,
                                         double (&m)[6]) {
    const Vec3 X0 = g.X[0], X1 = g.X[1], X2 = g.X[2], X3 = g.X[3];
    const double H00 = dot(X0, X0), H01 = dot(X0, X1), H11 = dot(X1, X1);
    const double H20 = dot(X2, X0), H21 = dot(X2, X1), H30 = dot(X3, X0), H31 = dot(X3, X1);
    m[0] = qd * (H11 + H00) - (qo + qo) * H01;
    m[1] = qd * H21 - qo * H20 - qo * H01 + qo * H00;
    m[2] = qd * H31 - qo * H30 - qo * H01 + qo * H00;
    m[3] = qo * H21 - qd * H20 - qo * H11 + qo * H01;
    m[4] = qo * H31 - qd * H30 - qo * H11 + qo * H01;
    m[5] = qo * (H31 - H30 - H21 + H20);
}

__device__ __forceinline__ void mass_col_f32(const FrameGeom& g, float qd, float qo,
                                             float (&m)[6]) {
    const Vec3f X0 = tof(g.X[0]), X1 = tof(g.X[1]), X2 = tof(g.X[2]), X3 = tof(g.X[3]);
    const float H00 = dotf(X0, X0), H01 = dotf(X0, X1), H11 = dotf(X1, X1);
    const float H20 = dotf(X2, X0), H21 = dotf(X2, X1), H30 = dotf(X3, X0), H31 = dotf(X3, X1);
    m[0] = qd * (H11 + H00) - (qo + qo) * H01;
    m[1] = qd * H21 - qo * H20 - qo * H01 + qo * H00;
    m[2] = qd * H31 - qo * H30 - qo * H01 + qo * H00;
    m[3] = qo * H21 - qd * H20 - qo * H11 + qo * H01;
    m[4] = qo * H31 - qd * H30 - qo * H11 + qo * H01;
    m[5] = qo * (H31 - H30 - H21 + H20);
}

// Column 0 of the curl-curl Gram (u_r' . u_0) in FP64 or FP32.
template <int kPrec>
__device__ __forceinline__ double curl_dot(const FrameGeom& g, int rp) {
    if (kPrec == MPK_PREC_FP32_LOCAL) return (double)dotf(tof(g.u[rp]), tof(g.u[0]));
    return dot(g.u[rp], g.u[0]);
}

template <int kPrec>
__device__ __forceinline__ void mass_col_prec(const FrameGeom& g, const CscQuad& q,
                                              double (&m)[6]) {
    if (kPrec == MPK_PREC_FP32_LOCAL) {
        float mf[6];
        mass_col_f32(g, (float)q.qd, (float)q.qo, mf);
#pragma unroll
        for (int r = 0; r < 6; ++r) m[r] = (double)mf[r];
    } else {
        mass_col(g, q.qd, q.qo, m);
    }
}

// Frame driver.  Per lane, a 4-stage pipeline over the column's incidences:
//   stage 3: frame record                           for k + 3
//   stage 2: element node ids, per-element data     for k + 2
//   stage 1: coordinates (+ op-specific gathers)    for k + 1
//   stage 0: arithmetic and LDS accumulation        for k
template <class Op>
__global__ void __launch_bounds__(kW) k_csc_frame(TileCtx tc, Op op) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Acc = typename Op::Acc;
    using Ids = typename Op::Ids;
    using Pre = typename Op::Pre;
    Acc* acc = reinterpret_cast<Acc*>(smem);
    const int lane = threadIdx.x;
    const int64_t tile_id = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t j0 = tile_id * kW;
    const int64_t jend = j0 + kW < tc.n ? j0 + kW : tc.n;
    const int base = tc.colptr[j0];
    const int tile = tc.colptr[jend] - base;
    op.block_init(smem + (size_t)op.acc_bytes_max);
    for (int i = lane; i < tile; i += kW) acc[i] = Op::zero();
    __syncthreads();
    const int64_t j = j0 + lane;
    if (j < jend) {
        const int cbase = tc.colptr[j];
        Acc* col = acc + (cbase - base);
        typename Op::Col cs = op.col_begin();
        const int kb = tc.inc_ptr[j], ke = tc.inc_ptr[j + 1];
        const int cnt = ke - kb;
        if (cnt > 0) {
            // k-major records: incidence k of this column at rec[k * kW]
            const uint4* __restrict__ rec = tc.inc_frame + tc.frame_base[tile_id] + lane;
            const uint4 z4 = make_uint4(0, 0, 0, 0);
            uint4 r0 = rec[0], r1 = z4, r2 = z4;
            if (1 < cnt) r1 = rec[kW];
            if (2 < cnt) r2 = rec[2 * kW];
            int4 t0 = tc.T[frec_elem(r0)], t1 = t0;
            Ids d0 = op.ids(tc, cbase, r0), d1 = d0;
            if (1 < cnt) {
                t1 = tc.T[frec_elem(r1)];
                d1 = op.ids(tc, cbase, r1);
            }
            // the column edge's endpoints, common to all its incidences
            const int2 ends = edge_ends(t0, frec_c(r0));
            const Vec3 xlo = ld_node(tc.P4, ends.x), xhi = ld_node(tc.P4, ends.y);
            Pre pre0 = op.fetch(tc, r0, t0, d0);
            for (int k = 0; k < cnt; ++k) {
                Pre pre1 = pre0;
                if (k + 1 < cnt) pre1 = op.fetch(tc, r1, t1, d1);
                int4 t2 = t1;
                Ids d2 = d1;
                if (k + 2 < cnt) {
                    t2 = tc.T[frec_elem(r2)];
                    d2 = op.ids(tc, cbase, r2);
                }
                uint4 r3 = z4;
                if (k + 3 < cnt) r3 = rec[(k + 3) * kW];
                op.consume(pre0, r0, xlo, xhi, col, cs);
                pre0 = pre1;
                r0 = r1; r1 = r2; r2 = r3;
                t1 = t2; d1 = d2;
            }
        }
        op.col_end(j, cs);
    }
    __syncthreads();
    for (int i = lane; i < tile; i += kW) op.store(base + i, acc[i]);
}

struct NoIds {};

// ---- frame ops ----
template <int kPrec>
struct FrameCurl {
    using Acc = double;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        FrameOpp opp;
        double nu;
    };
    const double* __restrict__ nu;
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { out[p] = a; }
    __device__ Ids ids(const TileCtx&, int, uint4) const { return Ids{}; }
    __device__ Pre fetch(const TileCtx& tc, uint4 r, int4 t, const Ids&) const {
        return Pre{fetch_opp(tc.P4, t, frec_c(r)), nu[frec_elem(r)]};
    }
    __device__ void consume(const Pre& pre, uint4 r, const Vec3& xlo, const Vec3& xhi, Acc* acc,
                            Col&) const {
        const FrameGeom g = frame_geom(frame_nodes(xlo, xhi, pre.opp), false);
        const double scale = 4.0 * qa.wsum * pre.nu / fabs(g.det) * frec_f(r, 0);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = frec_pos(r, rp);
            const double d = curl_dot<kPrec>(g, rp);
            if (p != 0xff) acc[p] += d * (scale * frec_f(r, rp));
        }
    }
};

template <int kPrec>
struct FrameMass {
    using Acc = double;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        FrameOpp opp;
        double sigma;
    };
    const double* __restrict__ coeffs;  // NULL: sigma = 1
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { out[p] = a; }
    __device__ Ids ids(const TileCtx&, int, uint4) const { return Ids{}; }
    __device__ Pre fetch(const TileCtx& tc, uint4 r, int4 t, const Ids&) const {
        return Pre{fetch_opp(tc.P4, t, frec_c(r)), coeffs ? coeffs[frec_elem(r)] : 1.0};
    }
    __device__ void consume(const Pre& pre, uint4 r, const Vec3& xlo, const Vec3& xhi, Acc* acc,
                            Col&) const {
        if (!(fabs(pre.sigma) > 1e-12)) return;  // zero block (mass_kernel_mex.cpp:69)
        const FrameGeom g = frame_geom(frame_nodes(xlo, xhi, pre.opp), true);
        const double scale = pre.sigma / fabs(g.det) * frec_f(r, 0);
        double m[6];
        mass_col_prec<kPrec>(g, qa, m);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = frec_pos(r, rp);
            if (p != 0xff) acc[p] += m[rp] * (scale * frec_f(r, rp));
        }
    }
};

template <int kPrec>
struct FrameKM {
    using Acc = double2;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        FrameOpp opp;
        double nu, sigma;
    };
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
        out[p] = make_double2(a.x, omega * a.y);
    }
    __device__ Ids ids(const TileCtx&, int, uint4) const { return Ids{}; }
    __device__ Pre fetch(const TileCtx& tc, uint4 r, int4 t, const Ids&) const {
        const uint32_t e = frec_elem(r);
        return Pre{fetch_opp(tc.P4, t, frec_c(r)), nu[e], sigma[e]};
    }
    __device__ void consume(const Pre& pre, uint4 r, const Vec3& xlo, const Vec3& xhi, Acc* acc,
                            Col&) const {
        const bool has_m = fabs(pre.sigma) > 1e-12;
        const FrameGeom g = frame_geom(frame_nodes(xlo, xhi, pre.opp), has_m);
        const double inv = 1.0 / fabs(g.det);
        const double f0 = frec_f(r, 0);
        const double ks = 4.0 * qk.wsum * pre.nu * inv * f0;
        const double ms = pre.sigma * inv * f0;
        double m[6] = {0, 0, 0, 0, 0, 0};
        if (has_m) mass_col_prec<kPrec>(g, qm, m);
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = frec_pos(r, rp);
            const double d = curl_dot<kPrec>(g, rp);
            const double f = frec_f(r, rp);
            if (p != 0xff) {
                Acc a = acc[p];
                a.x += d * (ks * f);
                if (has_m) a.y += m[rp] * (ms * f);
                acc[p] = a;
            }
        }
    }
};

// Newton Jacobian + residual in the frame.  The element's 6 edge DOFs are the
// column's own row indices (rowind[colptr[j] + pos'_r'] == dof - 1), read one
// stage before the coordinates; the solution values are read with them.
struct FrameRows {
    int row[6];  // 0-based DOF of frame edge r', -1 when inactive
};

template <int kPrec, bool kCalcJ>
struct FrameJac {
    using Acc = double;
    using Ids = FrameRows;
    struct Col {
        double r;
    };
    struct Pre {
        FrameOpp opp;
        double a[6];  // SolA of the frame edges (0 when inactive)
        int tag;
    };
    const int32_t* __restrict__ tags;
    const double* __restrict__ sol;
    MaterialView mat;
    int ntags, nbrk, ncoe;
    double* __restrict__ out;
    double* __restrict__ res;
    CscQuad qa;
    int acc_bytes_max;
    const double* brk_s;  // set per block (LDS or HBM)
    const double* coe_s;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char* lds) {
        if (nbrk + ncoe <= kMatLds) {
            double* tab = reinterpret_cast<double*>(lds);
            for (int i = threadIdx.x; i < nbrk; i += kW) tab[i] = mat.brk[i];
            for (int i = threadIdx.x; i < ncoe; i += kW) tab[nbrk + i] = mat.coe[i];
            brk_s = tab;
            coe_s = tab + nbrk;
        } else {
            brk_s = mat.brk;
            coe_s = mat.coe;
        }
    }
    __device__ Col col_begin() const { return Col{0.0}; }
    __device__ void col_end(int64_t j, Col& cs) const { res[j] = cs.r; }
    __device__ void store(int64_t p, Acc a) const {
        if (kCalcJ) out[p] = a;
    }
    __device__ Ids ids(const TileCtx& tc, int cbase, uint4 r) const {
        Ids d;
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const int p = frec_pos(r, rp);
            d.row[rp] = (p != 0xff) ? tc.rowind[cbase + p] : -1;
        }
        return d;
    }
    __device__ Pre fetch(const TileCtx& tc, uint4 r, int4 t, const Ids& d) const {
        Pre pre;
        pre.opp = fetch_opp(tc.P4, t, frec_c(r));
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) pre.a[rp] = d.row[rp] >= 0 ? sol[d.row[rp]] : 0.0;
        const int tag = tags[frec_elem(r)];
        // tags are validated by the host layer; clamp so a bad tag cannot read
        // outside the material table
        pre.tag = tag < 1 ? 1 : (tag > ntags ? ntags : tag);
        return pre;
    }
    __device__ void consume(const Pre& pre, uint4 r, const Vec3& xlo, const Vec3& xhi, Acc* acc,
                            Col& cs) const {
        const FrameGeom g = frame_geom(frame_nodes(xlo, xhi, pre.opp), false);
        // C'_r' = 2 u_r' / det;  B = sum a'_r' C'_r' with a' = A * f
        const double two_inv = 2.0 / g.det;
        Vec3 Bv{0.0, 0.0, 0.0};
#pragma unroll
        for (int rp = 0; rp < 6; ++rp) {
            const double ap = pre.a[rp] * frec_f(r, rp);
            Bv.x += ap * g.u[rp].x;
            Bv.y += ap * g.u[rp].y;
            Bv.z += ap * g.u[rp].z;
        }
        Bv.x *= two_inv;
        Bv.y *= two_inv;
        Bv.z *= two_inv;
        double nu, dnu;
        eval_nu(dot(Bv, Bv), pre.tag, mat, brk_s, coe_s, nu, dnu);
        // w|det| summed over the rule; C'.C' = 4 u.u / det^2
        const double wd = qa.wsum * fabs(g.det);
        const double f0 = frec_f(r, 0);
        const double BC0 = two_inv * dot(Bv, g.u[0]);
        cs.r += nu * BC0 * wd * f0;
        if (kCalcJ) {
            const double cc = 4.0 / (g.det * g.det);
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = frec_pos(r, rp);
                const double BCr = two_inv * dot(Bv, g.u[rp]);
                const double uu = curl_dot<kPrec>(g, rp);
                const double val = nu * (cc * uu) + 2.0 * dnu * BCr * BC0;
                if (p != 0xff) acc[p] += val * wd * (frec_f(r, rp) * f0);
            }
        }
    }
};

// =========================================================================
// Exact path (reference operation order, original node order)
// =========================================================================
struct NodesE {
    double p[12];
};

__device__ __forceinline__ NodesE fetch_orig(const double* __restrict__ P, int4 t) {
    NodesE g;
    const int nd[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        g.p[3 * n + 0] = P[3 * (int64_t)nd[n] + 0];
        g.p[3 * n + 1] = P[3 * (int64_t)nd[n] + 1];
        g.p[3 * n + 2] = P[3 * (int64_t)nd[n] + 2];
    }
    return g;
}

struct ExactGeom {
    double J[9];
    double det, absdet, invdet;
};

// compute_jacobian_3d, src/mex/MexElemUtils.hpp:35-49.
__device__ __forceinline__ ExactGeom exact_geom(const NodesE& ne) {
#pragma clang fp contract(off)
    ExactGeom g;
    double* J = g.J;
    const double* p = ne.p;
    J[0] = p[3] - p[0]; J[3] = p[4] - p[1]; J[6] = p[5] - p[2];
    J[1] = p[6] - p[0]; J[4] = p[7] - p[1]; J[7] = p[8] - p[2];
    J[2] = p[9] - p[0]; J[5] = p[10] - p[1]; J[8] = p[11] - p[2];
    g.det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
            J[2] * (J[3] * J[7] - J[4] * J[6]);
    g.absdet = fabs(g.det);
    g.invdet = 1.0 / g.det;
    return g;
}

// Mesh.T2E_Sign (src/Mesh.m:134-135): +1 iff T(a) <= T(b).
__device__ __forceinline__ void orig_signs(int4 t, double (&s)[6]) {
    const int nd[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = nd[edge_a(k)] <= nd[edge_b(k)] ? 1.0 : -1.0;
}

// K[r][c] for r = 0..5 (reference loop, assemble_curl_curl_kernel_mex.cpp:77-101).
__device__ __forceinline__ void exact_curl_col(const ExactGeom& g, int c, double nu,
                                               const CscQuad& qa, double (&k)[6]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < 6; ++r) k[r] = 0.0;
    for (int q = 0; q < qa.nq; ++q) {
        const double w = qa.qw[q] * g.absdet;
        double B[3][6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double r0 = qa.ref[i + 0 + 18 * q], r1 = qa.ref[i + 6 + 18 * q],
                         r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                B[a][i] = (g.J[3 * a + 0] * r0 + g.J[3 * a + 1] * r1 + g.J[3 * a + 2] * r2) * g.invdet;
        }
        const double c0 = sel6(c, B[0][0], B[0][1], B[0][2], B[0][3], B[0][4], B[0][5]);
        const double c1 = sel6(c, B[1][0], B[1][1], B[1][2], B[1][3], B[1][4], B[1][5]);
        const double c2 = sel6(c, B[2][0], B[2][1], B[2][2], B[2][3], B[2][4], B[2][5]);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double d = B[0][r] * c0 + B[1][r] * c1 + B[2][r] * c2;
            k[r] += d * nu * w;
        }
    }
}

// M[r][c] (assemble_mass_kernel_mex.cpp:63-108), sigma already tested > 1e-12.
__device__ __forceinline__ void exact_mass_col(const ExactGeom& g, int c, double sigma,
                                               const CscQuad& qa, double (&m)[6]) {
#pragma clang fp contract(off)
    const double* J = g.J;
    double Ji[9];
    const double invDet = 1.0 / g.det;
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) * invDet;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * invDet;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * invDet;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * invDet;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * invDet;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * invDet;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * invDet;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * invDet;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * invDet;
#pragma unroll
    for (int r = 0; r < 6; ++r) m[r] = 0.0;
    for (int q = 0; q < qa.nq; ++q) {
        const double w = qa.qw[q] * g.absdet;
        double N[3][6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double r0 = qa.ref[i + 0 + 18 * q], r1 = qa.ref[i + 6 + 18 * q],
                         r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
            for (int a = 0; a < 3; ++a) N[a][i] = Ji[0 + a] * r0 + Ji[3 + a] * r1 + Ji[6 + a] * r2;
        }
        const double c0 = sel6(c, N[0][0], N[0][1], N[0][2], N[0][3], N[0][4], N[0][5]);
        const double c1 = sel6(c, N[1][0], N[1][1], N[1][2], N[1][3], N[1][4], N[1][5]);
        const double c2 = sel6(c, N[2][0], N[2][1], N[2][2], N[2][3], N[2][4], N[2][5]);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double d = N[0][r] * c0 + N[1][r] * c1 + N[2][r] * c2;
            m[r] += d * sigma * w;
        }
    }
}

// Direct driver (exact mode): 4-stage pipeline over (slot, position word)
// records:
//   stage 3: incidence record                        for k + 3
//   stage 2: element node ids T[e] (+ op ids)        for k + 2
//   stage 1: coordinates and per-element data        for k + 1
//   stage 0: arithmetic and LDS accumulation         for k
template <class Op>
__global__ void __launch_bounds__(kW) k_csc_exact(TileCtx tc, Op op) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Acc = typename Op::Acc;
    using Pre = typename Op::Pre;
    using Ids = typename Op::Ids;
    Acc* acc = reinterpret_cast<Acc*>(smem);
    const int lane = threadIdx.x;
    const int64_t tile_id = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t j0 = tile_id * kW;
    const int64_t jend = j0 + kW < tc.n ? j0 + kW : tc.n;
    const int base = tc.colptr[j0];
    const int tile = tc.colptr[jend] - base;
    op.block_init(smem + (size_t)op.acc_bytes_max);
    for (int i = lane; i < tile; i += kW) acc[i] = Op::zero();
    __syncthreads();
    const int64_t j = j0 + lane;
    if (j < jend) {
        const int cbase = tc.colptr[j];
        Acc* col = acc + (cbase - base);
        typename Op::Col cs = op.col_begin();
        const int kb = tc.inc_ptr[j], ke = tc.inc_ptr[j + 1];
        if (kb < ke) {
            uint32_t s0 = tc.inc_slot[kb], s1 = 0, s2 = 0;
            uint2 p0 = tc.inc_pos[kb], p1 = make_uint2(0, 0), p2 = make_uint2(0, 0);
            if (kb + 1 < ke) { s1 = tc.inc_slot[kb + 1]; p1 = tc.inc_pos[kb + 1]; }
            if (kb + 2 < ke) { s2 = tc.inc_slot[kb + 2]; p2 = tc.inc_pos[kb + 2]; }
            int4 t0 = tc.T[s0 / 6u], t1 = t0;
            Ids i0 = op.ids(tc, cbase, p0), i1 = i0;
            if (kb + 1 < ke) { t1 = tc.T[s1 / 6u]; i1 = op.ids(tc, cbase, p1); }
            Pre pre0 = op.fetch(s0 / 6u, t0, i0);
            for (int k = kb; k < ke; ++k) {
                Pre pre1 = pre0;
                if (k + 1 < ke) pre1 = op.fetch(s1 / 6u, t1, i1);
                int4 t2 = t1;
                Ids i2 = i1;
                if (k + 2 < ke) { t2 = tc.T[s2 / 6u]; i2 = op.ids(tc, cbase, p2); }
                uint32_t s3 = 0;
                uint2 p3 = make_uint2(0, 0);
                if (k + 3 < ke) { s3 = tc.inc_slot[k + 3]; p3 = tc.inc_pos[k + 3]; }
                op.consume(pre0, (int)(s0 % 6u), t0, p0, col, cs);
                pre0 = pre1; t0 = t1; t1 = t2; i1 = i2;
                s0 = s1; s1 = s2; s2 = s3;
                p0 = p1; p1 = p2; p2 = p3;
            }
        }
        op.col_end(j, cs);
    }
    __syncthreads();
    for (int i = lane; i < tile; i += kW) op.store(base + i, acc[i]);
}

// ---- exact ops ----
struct ExactCurl {
    using Acc = double;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        NodesE nodes;
        double nu;
    };
    const double* __restrict__ P;
    const double* __restrict__ nu;
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { out[p] = a; }
    __device__ Ids ids(const TileCtx&, int, uint2) const { return Ids{}; }
    __device__ Pre fetch(uint32_t e, int4 t, const Ids&) const { return Pre{fetch_orig(P, t), nu[e]}; }
    __device__ void consume(const Pre& pre, int c, int4 t, uint2 pw, Acc* acc, Col&) const {
#pragma clang fp contract(off)
        const ExactGeom g = exact_geom(pre.nodes);
        double k[6], s[6];
        exact_curl_col(g, c, pre.nu, qa, k);
        orig_signs(t, s);
        const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int p = pos_of(pw, r);
            if (p != 0xff) acc[p] += k[r] * s[r] * sc;
        }
    }
};

struct ExactMass {
    using Acc = double;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        NodesE nodes;
        double sigma;
    };
    const double* __restrict__ P;
    const double* __restrict__ coeffs;  // NULL: sigma = 1
    double* __restrict__ out;
    CscQuad qa;
    int acc_bytes_max;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char*) const {}
    __device__ Col col_begin() const { return Col{}; }
    __device__ void col_end(int64_t, Col&) const {}
    __device__ void store(int64_t p, Acc a) const { out[p] = a; }
    __device__ Ids ids(const TileCtx&, int, uint2) const { return Ids{}; }
    __device__ Pre fetch(uint32_t e, int4 t, const Ids&) const {
        return Pre{fetch_orig(P, t), coeffs ? coeffs[e] : 1.0};
    }
    __device__ void consume(const Pre& pre, int c, int4 t, uint2 pw, Acc* acc, Col&) const {
#pragma clang fp contract(off)
        if (!(fabs(pre.sigma) > 1e-12)) return;  // zero block (mass_kernel_mex.cpp:69)
        const ExactGeom g = exact_geom(pre.nodes);
        double m[6], s[6];
        exact_mass_col(g, c, pre.sigma, qa, m);
        orig_signs(t, s);
        const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int p = pos_of(pw, r);
            if (p != 0xff) acc[p] += m[r] * s[r] * sc;
        }
    }
};

struct ExactKM {
    using Acc = double2;
    using Ids = NoIds;
    struct Col {};
    struct Pre {
        NodesE nodes;
        double nu, sigma;
    };
    const double* __restrict__ P;
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
        out[p] = make_double2(a.x, omega * a.y);
    }
    __device__ Ids ids(const TileCtx&, int, uint2) const { return Ids{}; }
    __device__ Pre fetch(uint32_t e, int4 t, const Ids&) const {
        return Pre{fetch_orig(P, t), nu[e], sigma[e]};
    }
    __device__ void consume(const Pre& pre, int c, int4 t, uint2 pw, Acc* acc, Col&) const {
#pragma clang fp contract(off)
        const bool has_m = fabs(pre.sigma) > 1e-12;
        const ExactGeom g = exact_geom(pre.nodes);
        double k[6], m[6], s[6];
        exact_curl_col(g, c, pre.nu, qk, k);
        if (has_m) exact_mass_col(g, c, pre.sigma, qm, m);
        orig_signs(t, s);
        const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int p = pos_of(pw, r);
            if (p != 0xff) {
                Acc a = acc[p];
                a.x += k[r] * s[r] * sc;
                if (has_m) a.y += m[r] * s[r] * sc;
                acc[p] = a;
            }
        }
    }
};

// The element's 6 edge DOFs are read from the column's own row indices
// (rowind[colptr[j] + pos_r] == dof_r - 1) one stage before the coordinates.
struct ExactRows {
    int row[6];  // 0-based DOF of original local edge r, -1 when inactive
};

template <bool kCalcJ>
struct ExactJac {
    using Acc = double;
    using Ids = ExactRows;
    struct Col {
        double r;
    };
    struct Pre {
        NodesE nodes;
        double a[6];  // SolA of the 6 original local edges (0 when inactive)
        int tag;
    };
    const double* __restrict__ P;
    const int32_t* __restrict__ tags;
    const double* __restrict__ sol;
    MaterialView mat;
    int ntags, nbrk, ncoe;
    double* __restrict__ out;
    double* __restrict__ res;
    CscQuad qa;
    int acc_bytes_max;
    const double* brk_s;  // set per block (LDS or HBM)
    const double* coe_s;
    __device__ static Acc zero() { return 0.0; }
    __device__ void block_init(unsigned char* lds) {
        if (nbrk + ncoe <= kMatLds) {
            double* tab = reinterpret_cast<double*>(lds);
            for (int i = threadIdx.x; i < nbrk; i += kW) tab[i] = mat.brk[i];
            for (int i = threadIdx.x; i < ncoe; i += kW) tab[nbrk + i] = mat.coe[i];
            brk_s = tab;
            coe_s = tab + nbrk;
        } else {
            brk_s = mat.brk;
            coe_s = mat.coe;
        }
    }
    __device__ Col col_begin() const { return Col{0.0}; }
    __device__ void col_end(int64_t j, Col& cs) const { res[j] = cs.r; }
    __device__ void store(int64_t p, Acc a) const {
        if (kCalcJ) out[p] = a;
    }
    __device__ Ids ids(const TileCtx& tc, int cbase, uint2 pw) const {
        Ids d;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int p = pos_of(pw, r);
            d.row[r] = (p != 0xff) ? tc.rowind[cbase + p] : -1;
        }
        return d;
    }
    __device__ Pre fetch(uint32_t e, int4 t, const Ids& d) const {
        Pre pre;
        pre.nodes = fetch_orig(P, t);
#pragma unroll
        for (int r = 0; r < 6; ++r) pre.a[r] = d.row[r] >= 0 ? sol[d.row[r]] : 0.0;
        const int tag = tags[e];
        pre.tag = tag < 1 ? 1 : (tag > ntags ? ntags : tag);
        return pre;
    }
    __device__ void consume(const Pre& pre, int c, int4 t, uint2 pw, Acc* acc, Col& cs) const {
#pragma clang fp contract(off)
        const int tag = pre.tag;
        const ExactGeom g = exact_geom(pre.nodes);
        double s[6], a[6];
        orig_signs(t, s);
        // A_local[i] = SolA[gid] * signs[i] (assemble_jacobian_kernel_mex.cpp:76-85)
#pragma unroll
        for (int i = 0; i < 6; ++i) a[i] = pre.a[i] * s[i];
        double kcol[6], rc = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) kcol[r] = 0.0;
        for (int q = 0; q < qa.nq; ++q) {
            const double w = qa.qw[q] * g.absdet;
            double C[6][3];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const double r0 = qa.ref[i + 0 + 18 * q], r1 = qa.ref[i + 6 + 18 * q],
                             r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    C[i][k] =
                        (g.J[3 * k + 0] * r0 + g.J[3 * k + 1] * r1 + g.J[3 * k + 2] * r2) * g.invdet;
            }
            double B0 = 0.0, B1 = 0.0, B2 = 0.0;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                B0 += a[i] * C[i][0];
                B1 += a[i] * C[i][1];
                B2 += a[i] * C[i][2];
            }
            const double Bsq = B0 * B0 + B1 * B1 + B2 * B2;
            double nu, dnu;
            eval_nu(Bsq, tag, mat, brk_s, coe_s, nu, dnu);
            const double cc0 = sel6(c, C[0][0], C[1][0], C[2][0], C[3][0], C[4][0], C[5][0]);
            const double cc1 = sel6(c, C[0][1], C[1][1], C[2][1], C[3][1], C[4][1], C[5][1]);
            const double cc2 = sel6(c, C[0][2], C[1][2], C[2][2], C[3][2], C[4][2], C[5][2]);
            const double BCc = B0 * cc0 + B1 * cc1 + B2 * cc2;
            rc += nu * BCc * w;
            if (kCalcJ) {
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    const double BCr = B0 * C[r][0] + B1 * C[r][1] + B2 * C[r][2];
                    const double CrCc = C[r][0] * cc0 + C[r][1] * cc1 + C[r][2] * cc2;
                    const double val = nu * CrCc + 2.0 * dnu * BCr * BCc;
                    kcol[r] += val * w;
                }
            }
        }
        const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
        cs.r += rc * sc;
        if (kCalcJ) {
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                const int p = pos_of(pw, r);
                if (p != 0xff) acc[p] += kcol[r] * s[r] * sc;
            }
        }
    }
};

// ------------------------------------------------------------------------
// Host helpers
// ------------------------------------------------------------------------
__global__ void k_pad_coords(const double* __restrict__ P, int64_t nn, double4* __restrict__ P4) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    P4[i] = make_double4(P[3 * i], P[3 * i + 1], P[3 * i + 2], 0.0);
}

int make_csc_quad(const mpk_quadrature* q, bool curl, CscQuad& a) {
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

// The closed forms assume the Nedelec P1 reference curls of
// nedelec_tet_p1.m:48 at every point (constant curl).
bool standard_curl_table(const mpk_quadrature* q) {
    static const double c[6][3] = {{0, -2, 2}, {2, 0, -2}, {-2, 2, 0},
                                   {0, 0, 2},  {0, -2, 0}, {2, 0, 0}};
    for (int k = 0; k < q->nq; ++k)
        for (int i = 0; i < 6; ++i)
            for (int a = 0; a < 3; ++a)
                if (q->ref_curl[i + 6 * a + 18 * k] != c[i][a]) return false;
    return true;
}

TileCtx tile_ctx(const mpk_plan* plan) {
    return TileCtx{plan->colptr,     plan->inc_ptr, plan->inc_slot, plan->inc_pos,
                   plan->inc_frame,  plan->frame_base, plan->T,   plan->P4,
                   plan->rowind,     plan->n};
}

template <template <class> class K, class Op>
int launch(const mpk_plan* plan, Op op, void (*kernel)(TileCtx, Op), size_t acc_size,
           size_t extra_lds, hipStream_t st) {
    if (plan->n == 0) return MPK_OK;
    const size_t acc_bytes = ((size_t)plan->max_tile_nnz * acc_size + 15) & ~size_t(15);
    const size_t lds = acc_bytes + extra_lds;
    MPK_REQUIRE(lds <= (size_t)kMaxLdsBytes, MPK_ERR_LIMIT,
                "tile of %lld entries needs %zu B of LDS (limit %d)",
                (long long)plan->max_tile_nnz, lds, kMaxLdsBytes);
    op.acc_bytes_max = (int)acc_bytes;
    hipLaunchKernelGGL(kernel, dim3(grid_for(plan->n, kW)), dim3(kW), lds, st, tile_ctx(plan), op);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

template <class Op>
struct ExactK {};
template <class Op>
struct FrameK {};

template <class Op>
int launch_exact(const mpk_plan* plan, Op op, size_t acc_size, size_t extra, hipStream_t st) {
    return launch<ExactK>(plan, op, &k_csc_exact<Op>, acc_size, extra, st);
}

// Frame kernels read the padded coordinate copy: refresh it from P first
// (stream-ordered, so a P changed by the caller between calls is picked up).
template <class Op>
int launch_frame(const mpk_plan* plan, const double* P, Op op, size_t acc_size, size_t extra,
                 hipStream_t st) {
    if (plan->n == 0) return MPK_OK;
    if (plan->nn > 0 && plan->P4) {
        hipLaunchKernelGGL(k_pad_coords, dim3(grid_for(plan->nn, 256)), dim3(256), 0, st, P,
                           plan->nn, plan->P4);
        MPK_HIP_TRY(hipGetLastError());
    }
    return launch<FrameK>(plan, op, &k_csc_frame<Op>, acc_size, extra, st);
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_csc_curl_curl(const mpk_plan* plan, const double* P, const double* nu,
                      const mpk_quadrature* quad, int precision, double* vals, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(plan->nnz == 0 || (P && nu && vals), MPK_ERR_ARGS, "NULL array argument");
    CscQuad qa;
    if (int rc = make_csc_quad(quad, true, qa)) return rc;
    if (precision != MPK_PREC_FP64_EXACT && !standard_curl_table(quad))
        precision = MPK_PREC_FP64_EXACT;
    hipStream_t st = as_stream(stream);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            return launch_exact(plan, ExactCurl{P, nu, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP64:
            return launch_frame(plan, P, FrameCurl<MPK_PREC_FP64>{nu, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_frame(plan, P, FrameCurl<MPK_PREC_FP32_LOCAL>{nu, vals, qa, 0}, 8, 0, st);
    }
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

int mpk_csc_mass(const mpk_plan* plan, const double* P, const double* coeffs,
                 const mpk_quadrature* quad, int precision, double* vals, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(plan->nnz == 0 || (P && vals), MPK_ERR_ARGS, "NULL array argument");
    CscQuad qa;
    if (int rc = make_csc_quad(quad, false, qa)) return rc;
    hipStream_t st = as_stream(stream);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            return launch_exact(plan, ExactMass{P, coeffs, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP64:
            return launch_frame(plan, P, FrameMass<MPK_PREC_FP64>{coeffs, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_frame(plan, P, FrameMass<MPK_PREC_FP32_LOCAL>{coeffs, vals, qa, 0}, 8,
                                0, st);
    }
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

int mpk_csc_curl_curl_mass_complex(const mpk_plan* plan, const double* P, const double* nu,
                                   const double* sigma, double omega,
                                   const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                                   int precision, double* vals_c128, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(plan->nnz == 0 || (P && nu && sigma && vals_c128), MPK_ERR_ARGS,
                "NULL array argument");
    CscQuad qk, qm;
    if (int rc = make_csc_quad(quad_k, true, qk)) return rc;
    if (int rc = make_csc_quad(quad_m, false, qm)) return rc;
    if (precision != MPK_PREC_FP64_EXACT && !standard_curl_table(quad_k))
        precision = MPK_PREC_FP64_EXACT;
    hipStream_t st = as_stream(stream);
    double2* out = reinterpret_cast<double2*>(vals_c128);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            return launch_exact(plan, ExactKM{P, nu, sigma, out, omega, qk, qm, 0}, 16, 0, st);
        case MPK_PREC_FP64:
            return launch_frame(plan, P, FrameKM<MPK_PREC_FP64>{nu, sigma, out, omega, qk, qm, 0},
                                16, 0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_frame(plan, P,
                                FrameKM<MPK_PREC_FP32_LOCAL>{nu, sigma, out, omega, qk, qm, 0},
                                16, 0, st);
    }
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

int mpk_csc_jacobian(const mpk_plan* plan, const double* P, const int32_t* tags,
                     const double* sol_a, const mpk_material* mat,
                     const mpk_quadrature* quad, int calc_j, int precision,
                     double* j_vals, double* residual, void* stream) {
    clear_error();
    MPK_REQUIRE(plan && mat, MPK_ERR_ARGS, "NULL plan or material");
    MPK_REQUIRE(plan->n == 0 || (P && tags && sol_a && residual), MPK_ERR_ARGS,
                "NULL array argument");
    MPK_REQUIRE(!calc_j || plan->nnz == 0 || j_vals, MPK_ERR_ARGS, "NULL j_vals");
    CscQuad qa;
    if (int rc = make_csc_quad(quad, true, qa)) return rc;
    if (precision != MPK_PREC_FP64_EXACT && !standard_curl_table(quad))
        precision = MPK_PREC_FP64_EXACT;
    MaterialView mv{mat->linear_nu, mat->is_nonlinear, mat->max_bsq,    mat->breaks,
                    mat->coefs,     mat->spline_start, mat->spline_count, mat->coef_start};
    const int nbrk = (int)mat->num_breaks, ncoe = (int)mat->num_coefs;
    const int ntags = (int)mat->num_tags;
    MPK_REQUIRE(ntags >= 1, MPK_ERR_ARGS, "material table without tags");
    const size_t extra = (nbrk + ncoe <= kMatLds) ? (size_t)(nbrk + ncoe) * 8 : 0;
    hipStream_t st = as_stream(stream);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            if (calc_j)
                return launch_exact(plan, ExactJac<true>{P, tags, sol_a, mv, ntags, nbrk, ncoe,
                                                         j_vals, residual, qa, 0, nullptr, nullptr},
                                    8, extra, st);
            return launch_exact(plan, ExactJac<false>{P, tags, sol_a, mv, ntags, nbrk, ncoe,
                                                      j_vals, residual, qa, 0, nullptr, nullptr},
                                8, extra, st);
        case MPK_PREC_FP64:
            if (calc_j)
                return launch_frame(plan, P,
                                    FrameJac<MPK_PREC_FP64, true>{tags, sol_a, mv, ntags, nbrk,
                                                                  ncoe, j_vals, residual, qa, 0,
                                                                  nullptr, nullptr},
                                    8, extra, st);
            return launch_frame(plan, P,
                                FrameJac<MPK_PREC_FP64, false>{tags, sol_a, mv, ntags, nbrk, ncoe,
                                                               j_vals, residual, qa, 0, nullptr,
                                                               nullptr},
                                8, extra, st);
        case MPK_PREC_FP32_LOCAL:
            if (calc_j)
                return launch_frame(plan, P,
                                    FrameJac<MPK_PREC_FP32_LOCAL, true>{tags, sol_a, mv, ntags,
                                                                        nbrk, ncoe, j_vals,
                                                                        residual, qa, 0, nullptr,
                                                                        nullptr},
                                    8, extra, st);
            return launch_frame(plan, P,
                                FrameJac<MPK_PREC_FP32_LOCAL, false>{tags, sol_a, mv, ntags, nbrk,
                                                                     ncoe, j_vals, residual, qa, 0,
                                                                     nullptr, nullptr},
                                8, extra, st);
    }
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

}  // extern "C"

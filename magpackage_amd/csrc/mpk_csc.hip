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
// Element arithmetic, per precision mode:
//   FP64_EXACT  reference operation order, no FMA contraction:
//               src/mex/assemble_curl_curl_kernel_mex.cpp:57-101,
//               src/mex/assemble_mass_kernel_mex.cpp:58-108,
//               src/mex/assemble_jacobian_kernel_mex.cpp:71-131.
//   FP64        closed forms in a frame where the lane's edge is local edge
//               0 (nodes permuted), FMA allowed.  Curl-curl:
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
    const int4* __restrict__ T;
    const int32_t* __restrict__ rowind;
    int64_t n;
};

__device__ __forceinline__ int pos_of(uint2 pw, int r) {
    return r < 4 ? (int)((pw.x >> (8 * r)) & 0xffu) : (int)((pw.y >> (8 * (r - 4))) & 0xffu);
}

// Packed map (permuted local edge r') -> original local edge, 3 bits each,
// for the node permutation that makes edge c the permuted edge (0,1):
//   c=0 (0,1,2,3)  c=1 (0,2,1,3)  c=2 (0,3,1,2)  c=3 (1,2,0,3)  c=4 (1,3,0,2)  c=5 (2,3,0,1)
__host__ __device__ constexpr int edge_index(int a, int b) {
    // a < b
    return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}
__host__ __device__ constexpr int perm_node(int c, int i) {
    // node permutation of row c, one nibble per position
    return (int)(((c == 0 ? 0x3210u : c == 1 ? 0x3120u : c == 2 ? 0x2130u
                 : c == 3 ? 0x3021u : c == 4 ? 0x2031u : 0x1032u) >> (4 * i)) & 0xfu);
}
__host__ __device__ constexpr int orig_edge(int c, int rp) {
    const int a = perm_node(c, edge_a(rp)), b = perm_node(c, edge_b(rp));
    return a < b ? edge_index(a, b) : edge_index(b, a);
}
__host__ __device__ constexpr uint32_t perm_word(int c) {
    uint32_t w = 0;
    for (int rp = 0; rp < 6; ++rp) w |= (uint32_t)orig_edge(c, rp) << (3 * rp);
    return w;
}
static_assert(orig_edge(0, 0) == 0 && orig_edge(5, 0) == 5 && orig_edge(5, 5) == 0 &&
                  orig_edge(2, 5) == 3 && orig_edge(4, 3) == 2,
              "permutation tables");

template <class T>
__device__ __forceinline__ T sel6(int c, T v0, T v1, T v2, T v3, T v4, T v5) {
    return c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : c == 3 ? v3 : c == 4 ? v4 : v5;
}

struct Vec3 {
    double x, y, z;
};
__device__ __forceinline__ Vec3 ld3(const double* __restrict__ P, int node) {
    return Vec3{P[3 * (int64_t)node], P[3 * (int64_t)node + 1], P[3 * (int64_t)node + 2]};
}
__device__ __forceinline__ Vec3 sub(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Geometry of one incidence in the permuted frame (lane edge == edge (0,1)).
struct FrameGeom {
    Vec3 u[6];      // opposite-edge vectors: J c_r' = 2 u_r'
    Vec3 X[4];      // det * grad lambda_i (rows of adj(J))
    double det;     // signed determinant in the permuted frame
    double f[6];    // orientation of permuted edge r' relative to the global edge
    uint32_t pword; // permuted -> original edge map
};

__device__ __forceinline__ FrameGeom frame_geom(const double* __restrict__ P, int4 t, int c,
                                                bool need_x) {
    const int n0 = sel6(c, t.x, t.x, t.x, t.y, t.y, t.z);
    const int n1 = sel6(c, t.y, t.z, t.w, t.z, t.w, t.w);
    const int n2 = sel6(c, t.z, t.y, t.y, t.x, t.x, t.x);
    const int n3 = sel6(c, t.w, t.w, t.z, t.w, t.z, t.y);
    FrameGeom g;
    const Vec3 x0 = ld3(P, n0), x1 = ld3(P, n1), x2 = ld3(P, n2), x3 = ld3(P, n3);
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
    const int nn[4] = {n0, n1, n2, n3};
#pragma unroll
    for (int rp = 0; rp < 6; ++rp) g.f[rp] = nn[edge_a(rp)] < nn[edge_b(rp)] ? 1.0 : -1.0;
    g.pword = sel6(c, perm_word(0), perm_word(1), perm_word(2), perm_word(3), perm_word(4),
                   perm_word(5));
    return g;
}

__device__ __forceinline__ int frame_pos(const FrameGeom& g, uint2 pw, int rp) {
    const int r = (int)((g.pword >> (3 * rp)) & 7u);
    return r < 4 ? (int)((pw.x >> (8 * r)) & 0xffu) : (int)((pw.y >> (8 * (r - 4))) & 0xffu);
}

// Mass column 0 in the permuted frame (without sigma/|det| and signs):
// sum over the rule of W_r' . W_0, with symmetric moments qd/qo.
template <class R>
__device__ __forceinline__ void mass_col(const FrameGeom& g, R qd, R qo, R (&m)[6]) {
    const Vec3 X0 = g.X[0], X1 = g.X[1], X2 = g.X[2], X3 = g.X[3];
    const R H00 = (R)dot(X0, X0), H01 = (R)dot(X0, X1), H11 = (R)dot(X1, X1);
    const R H20 = (R)dot(X2, X0), H21 = (R)dot(X2, X1), H30 = (R)dot(X3, X0), H31 = (R)dot(X3, X1);
    m[0] = qd * (H11 + H00) - (qo + qo) * H01;
    m[1] = qd * H21 - qo * H20 - qo * H01 + qo * H00;
    m[2] = qd * H31 - qo * H30 - qo * H01 + qo * H00;
    m[3] = qo * H21 - qd * H20 - qo * H11 + qo * H01;
    m[4] = qo * H31 - qd * H30 - qo * H11 + qo * H01;
    m[5] = qo * (H31 - H30 - H21 + H20);
}

struct Vec3f {
    float x, y, z;
};
__device__ __forceinline__ Vec3f tof(Vec3 v) { return Vec3f{(float)v.x, (float)v.y, (float)v.z}; }
__device__ __forceinline__ float dotf(Vec3f a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

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

// ------------------------------------------------------------------------
// Exact (reference-order) element columns, original node order.
// ------------------------------------------------------------------------
struct ExactGeom {
    double J[9];
    double det, absdet, invdet;
    double p[12];
};

__device__ __forceinline__ void exact_geom(const double* __restrict__ P, int4 t, ExactGeom& g) {
#pragma clang fp contract(off)
    const int nd[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        g.p[3 * n + 0] = P[3 * (int64_t)nd[n] + 0];
        g.p[3 * n + 1] = P[3 * (int64_t)nd[n] + 1];
        g.p[3 * n + 2] = P[3 * (int64_t)nd[n] + 2];
    }
    double* J = g.J;
    const double* p = g.p;
    J[0] = p[3] - p[0]; J[3] = p[4] - p[1]; J[6] = p[5] - p[2];
    J[1] = p[6] - p[0]; J[4] = p[7] - p[1]; J[7] = p[8] - p[2];
    J[2] = p[9] - p[0]; J[5] = p[10] - p[1]; J[8] = p[11] - p[2];
    g.det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
            J[2] * (J[3] * J[7] - J[4] * J[6]);
    g.absdet = fabs(g.det);
    g.invdet = 1.0 / g.det;
}

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

// ------------------------------------------------------------------------
// Tile driver
// ------------------------------------------------------------------------
template <class Op>
__global__ void __launch_bounds__(kW) k_csc_tile(TileCtx tc, Op op) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Acc = typename Op::Acc;
    Acc* acc = reinterpret_cast<Acc*>(smem);
    const int lane = threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.x * kW;
    const int64_t jend = j0 + kW < tc.n ? j0 + kW : tc.n;
    const int base = tc.colptr[j0];
    const int tile = tc.colptr[jend] - base;
    op.block_init(smem + (size_t)op.acc_bytes_max);
    for (int i = lane; i < tile; i += kW) acc[i] = Op::zero();
    __syncthreads();
    const int64_t j = j0 + lane;
    if (j < jend) {
        const int off = tc.colptr[j] - base;
        typename Op::Col cs = op.col_begin();
        const int kb = tc.inc_ptr[j], ke = tc.inc_ptr[j + 1];
        for (int k = kb; k < ke; ++k) {
            const uint32_t slot = tc.inc_slot[k];
            const uint2 pw = tc.inc_pos[k];
            const uint32_t e = slot / 6u;
            const int c = (int)(slot - 6u * e);
            op.incidence(tc, (int)e, c, pw, acc + off, tc.colptr[j], cs);
        }
        op.col_end(j, cs);
    }
    __syncthreads();
    for (int i = lane; i < tile; i += kW) op.store(base + i, acc[i]);
}

// ---------------------------- curl-curl -----------------------------------
template <int kPrec>
struct OpCurl {
    using Acc = double;
    struct Col {};
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
    __device__ void incidence(const TileCtx& tc, int e, int c, uint2 pw, Acc* acc, int,
                              Col&) const {
        const int4 t = tc.T[e];
        if (kPrec == MPK_PREC_FP64_EXACT) {
#pragma clang fp contract(off)
            ExactGeom g;
            exact_geom(P, t, g);
            double k[6], s[6];
            exact_curl_col(g, c, nu[e], qa, k);
            orig_signs(t, s);
            const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                const int p = pos_of(pw, r);
                if (p != 0xff) acc[p] += k[r] * s[r] * sc;
            }
        } else {
            const FrameGeom g = frame_geom(P, t, c, false);
            const double scale = 4.0 * qa.wsum * nu[e] / fabs(g.det) * g.f[0];
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = frame_pos(g, pw, rp);
                double d;
                if (kPrec == MPK_PREC_FP32_LOCAL)
                    d = (double)dotf(tof(g.u[rp]), tof(g.u[0]));
                else
                    d = dot(g.u[rp], g.u[0]);
                if (p != 0xff) acc[p] += d * (scale * g.f[rp]);
            }
        }
    }
};

// ------------------------------ mass --------------------------------------
template <int kPrec>
struct OpMass {
    using Acc = double;
    struct Col {};
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
    __device__ void incidence(const TileCtx& tc, int e, int c, uint2 pw, Acc* acc, int,
                              Col&) const {
        const double sigma = coeffs ? coeffs[e] : 1.0;
        if (!(fabs(sigma) > 1e-12)) return;  // zero block (mass_kernel_mex.cpp:69)
        const int4 t = tc.T[e];
        if (kPrec == MPK_PREC_FP64_EXACT) {
#pragma clang fp contract(off)
            ExactGeom g;
            exact_geom(P, t, g);
            double m[6], s[6];
            exact_mass_col(g, c, sigma, qa, m);
            orig_signs(t, s);
            const double sc = sel6(c, s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                const int p = pos_of(pw, r);
                if (p != 0xff) acc[p] += m[r] * s[r] * sc;
            }
        } else {
            const FrameGeom g = frame_geom(P, t, c, true);
            const double scale = sigma / fabs(g.det) * g.f[0];
            double m[6];
            if (kPrec == MPK_PREC_FP32_LOCAL) {
                float mf[6];
                mass_col_f32(g, (float)qa.qd, (float)qa.qo, mf);
#pragma unroll
                for (int r = 0; r < 6; ++r) m[r] = (double)mf[r];
            } else {
                mass_col<double>(g, qa.qd, qa.qo, m);
            }
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = frame_pos(g, pw, rp);
                if (p != 0xff) acc[p] += m[rp] * (scale * g.f[rp]);
            }
        }
    }
};

// ---------------------- K + j omega M_sigma (complex) -----------------------
template <int kPrec>
struct OpKM {
    using Acc = double2;
    struct Col {};
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
        // K + 1j*omega*M: real part K, imaginary part omega * (sum of M)
        out[p] = make_double2(a.x, omega * a.y);
    }
    __device__ void incidence(const TileCtx& tc, int e, int c, uint2 pw, Acc* acc, int,
                              Col&) const {
        const int4 t = tc.T[e];
        const double sg = sigma[e];
        const bool has_m = fabs(sg) > 1e-12;
        if (kPrec == MPK_PREC_FP64_EXACT) {
#pragma clang fp contract(off)
            ExactGeom g;
            exact_geom(P, t, g);
            double k[6], m[6], s[6];
            exact_curl_col(g, c, nu[e], qk, k);
            if (has_m) exact_mass_col(g, c, sg, qm, m);
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
        } else {
            const FrameGeom g = frame_geom(P, t, c, has_m);
            const double inv = 1.0 / fabs(g.det);
            const double ks = 4.0 * qk.wsum * nu[e] * inv * g.f[0];
            const double ms = sg * inv * g.f[0];
            double m[6];
            if (has_m) {
                if (kPrec == MPK_PREC_FP32_LOCAL) {
                    float mf[6];
                    mass_col_f32(g, (float)qm.qd, (float)qm.qo, mf);
#pragma unroll
                    for (int r = 0; r < 6; ++r) m[r] = (double)mf[r];
                } else {
                    mass_col<double>(g, qm.qd, qm.qo, m);
                }
            }
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = frame_pos(g, pw, rp);
                double d;
                if (kPrec == MPK_PREC_FP32_LOCAL)
                    d = (double)dotf(tof(g.u[rp]), tof(g.u[0]));
                else
                    d = dot(g.u[rp], g.u[0]);
                if (p != 0xff) {
                    Acc a = acc[p];
                    a.x += d * (ks * g.f[rp]);
                    if (has_m) a.y += m[rp] * (ms * g.f[rp]);
                    acc[p] = a;
                }
            }
        }
    }
};

// ------------------------- Newton Jacobian + R ------------------------------
template <int kPrec, bool kCalcJ>
struct OpJac {
    using Acc = double;
    struct Col {
        double r;
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
    __device__ void incidence(const TileCtx& tc, int e, int c, uint2 pw, Acc* acc, int cbase,
                              Col& cs) const {
        const int4 t = tc.T[e];
        int tag = tags[e];
        // tags are validated by the host layer; clamp so a bad tag cannot read
        // outside the material table
        tag = tag < 1 ? 1 : (tag > ntags ? ntags : tag);
        if (kPrec == MPK_PREC_FP64_EXACT) {
#pragma clang fp contract(off)
            ExactGeom g;
            exact_geom(P, t, g);
            double s[6], a[6];
            orig_signs(t, s);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const int p = pos_of(pw, i);
                a[i] = (p != 0xff) ? sol[tc.rowind[cbase + p]] * s[i] : 0.0;
            }
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
                        C[i][k] = (g.J[3 * k + 0] * r0 + g.J[3 * k + 1] * r1 + g.J[3 * k + 2] * r2) *
                                  g.invdet;
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
        } else {
            const FrameGeom g = frame_geom(P, t, c, false);
            // a'_r' = A(dof of the original edge) * f_r'
            double ap[6];
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                const int p = frame_pos(g, pw, rp);
                ap[rp] = (p != 0xff) ? sol[tc.rowind[cbase + p]] * g.f[rp] : 0.0;
            }
            // C'_r' = 2 u_r' / det;  B = sum a' C'
            const double two_inv = 2.0 / g.det;
            Vec3 Bv{0.0, 0.0, 0.0};
#pragma unroll
            for (int rp = 0; rp < 6; ++rp) {
                Bv.x += ap[rp] * g.u[rp].x;
                Bv.y += ap[rp] * g.u[rp].y;
                Bv.z += ap[rp] * g.u[rp].z;
            }
            Bv.x *= two_inv;
            Bv.y *= two_inv;
            Bv.z *= two_inv;
            const double Bsq = dot(Bv, Bv);
            double nu, dnu;
            eval_nu(Bsq, tag, mat, brk_s, coe_s, nu, dnu);
            // w|det| summed over the rule; C'.C' = 4 u.u / det^2
            const double wd = qa.wsum * fabs(g.det);
            const double BC0 = two_inv * dot(Bv, g.u[0]);
            cs.r += nu * BC0 * wd * g.f[0];
            if (kCalcJ) {
                const double cc = 4.0 / (g.det * g.det);
#pragma unroll
                for (int rp = 0; rp < 6; ++rp) {
                    const int p = frame_pos(g, pw, rp);
                    const double BCr = two_inv * dot(Bv, g.u[rp]);
                    double uu;
                    if (kPrec == MPK_PREC_FP32_LOCAL)
                        uu = (double)dotf(tof(g.u[rp]), tof(g.u[0]));
                    else
                        uu = dot(g.u[rp], g.u[0]);
                    const double val = nu * (cc * uu) + 2.0 * dnu * BCr * BC0;
                    if (p != 0xff) acc[p] += val * wd * (g.f[rp] * g.f[0]);
                }
            }
        }
    }
};

// ------------------------------------------------------------------------
// Host helpers
// ------------------------------------------------------------------------
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

template <class Op>
int launch_tile(const mpk_plan* plan, Op op, size_t acc_size, size_t extra_lds, hipStream_t st) {
    if (plan->n == 0) return MPK_OK;
    const size_t acc_bytes = ((size_t)plan->max_tile_nnz * acc_size + 15) & ~size_t(15);
    const size_t lds = acc_bytes + extra_lds;
    MPK_REQUIRE(lds <= (size_t)kMaxLdsBytes, MPK_ERR_LIMIT,
                "tile of %lld entries needs %zu B of LDS (limit %d)",
                (long long)plan->max_tile_nnz, lds, kMaxLdsBytes);
    op.acc_bytes_max = (int)acc_bytes;
    TileCtx tc{plan->colptr, plan->inc_ptr, plan->inc_slot, plan->inc_pos,
               plan->T,      plan->rowind,  plan->n};
    const unsigned grid = grid_for(plan->n, kW);
    hipLaunchKernelGGL(k_csc_tile<Op>, dim3(grid), dim3(kW), lds, st, tc, op);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
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
            return launch_tile(plan, OpCurl<MPK_PREC_FP64_EXACT>{P, nu, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP64:
            return launch_tile(plan, OpCurl<MPK_PREC_FP64>{P, nu, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_tile(plan, OpCurl<MPK_PREC_FP32_LOCAL>{P, nu, vals, qa, 0}, 8, 0, st);
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
            return launch_tile(plan, OpMass<MPK_PREC_FP64_EXACT>{P, coeffs, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP64:
            return launch_tile(plan, OpMass<MPK_PREC_FP64>{P, coeffs, vals, qa, 0}, 8, 0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_tile(plan, OpMass<MPK_PREC_FP32_LOCAL>{P, coeffs, vals, qa, 0}, 8, 0, st);
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
            return launch_tile(plan, OpKM<MPK_PREC_FP64_EXACT>{P, nu, sigma, out, omega, qk, qm, 0},
                               16, 0, st);
        case MPK_PREC_FP64:
            return launch_tile(plan, OpKM<MPK_PREC_FP64>{P, nu, sigma, out, omega, qk, qm, 0}, 16,
                               0, st);
        case MPK_PREC_FP32_LOCAL:
            return launch_tile(plan,
                               OpKM<MPK_PREC_FP32_LOCAL>{P, nu, sigma, out, omega, qk, qm, 0},
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
    auto go = [&](auto op) { return launch_tile(plan, op, 8, extra, st); };
#define MPK_JAC(PREC)                                                                          \
    (calc_j ? go(OpJac<PREC, true>{P, tags, sol_a, mv, ntags, nbrk, ncoe, j_vals, residual, \
                                   qa, 0, nullptr, nullptr})                                   \
            : go(OpJac<PREC, false>{P, tags, sol_a, mv, ntags, nbrk, ncoe, j_vals, residual,   \
                                    qa, 0, nullptr, nullptr}))
    switch (precision) {
        case MPK_PREC_FP64_EXACT: return MPK_JAC(MPK_PREC_FP64_EXACT);
        case MPK_PREC_FP64: return MPK_JAC(MPK_PREC_FP64);
        case MPK_PREC_FP32_LOCAL: return MPK_JAC(MPK_PREC_FP32_LOCAL);
    }
#undef MPK_JAC
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

}  // extern "C"

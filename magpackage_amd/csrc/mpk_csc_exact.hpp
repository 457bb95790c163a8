// Bit-exact CSC assembly (FP64_EXACT mode): the reference's operation order
// in the element's own node order, no FMA contraction, and every CSC entry
// summed in the triplet (element) order of the reference's COO output:
//   src/mex/assemble_curl_curl_kernel_mex.cpp:57-101,
//   src/mex/assemble_mass_kernel_mex.cpp:58-108,
//   src/mex/assemble_jacobian_kernel_mex.cpp:71-131.
#pragma once

#include "mpk_csc_common.hpp"

namespace mpk {
namespace csc {

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

struct ExactNoIds {};

// Direct driver (element order): 4-stage pipeline over (slot, position word)
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
    const TileRange tr = tile_begin(tc, op, acc, smem);
    const int64_t j = tr.j0 + threadIdx.x;
    if (j < tr.jend) {
        const int cbase = tc.colptr[j];
        Acc* col = acc + (cbase - tr.base);
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
    tile_end(op, tr, acc);
}

struct ExactCurl {
    using Acc = double;
    using Ids = ExactNoIds;
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
    using Ids = ExactNoIds;
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
    using Ids = ExactNoIds;
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
    __device__ void store(int64_t p, Acc a) const { out[p] = make_double2(a.x, omega * a.y); }
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
                double2 a = acc[p];
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

}  // namespace csc
}  // namespace mpk

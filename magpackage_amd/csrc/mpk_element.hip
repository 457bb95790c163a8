// MEX-compatible element kernels (COO triplets) on gfx950.
//
// One element per lane.  Each lane gathers its tetrahedron, forms the local
// 6x6 block in FP64 with the reference operation order (no FMA
// contraction), and the block of 64 lanes stages its 36*64 triplet values in
// LDS so that the I/J/V streams leave the CU as contiguous, coalesced stores
// in the reference's element-major order (slot 36*e + 6*c + r).
//
// Reference kernels restated:
//   k_curl_curl_coo  src/mex/assemble_curl_curl_kernel_mex.cpp:44-121
//   k_mass_coo       src/mex/assemble_mass_kernel_mex.cpp:46-127
//   k_jacobian_coo   src/mex/assemble_jacobian_kernel_mex.cpp:61-152
//   k_scalar_laplacian_coo  src/mex/assemble_scalar_laplacian_kernel_mex.cpp:41-104
//   k_coupling_coo   src/mex/assemble_coupling_kernel_mex.cpp:44-131
//   k_winding_coo    src/mex/assemble_winding_kernel_mex.cpp:48-125
//   k_hbfem_coo      src/mex/assemble_hbfem_kernel_mex.cpp:93-259
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "mpk_common.hpp"

namespace mpk {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

void clear_error() { g_err.clear(); }

namespace {

constexpr int kBlock = 64;         // one wave64 per block
constexpr int kStageStride = 37;   // 36 slots + 1 pad against LDS bank conflicts

// Quadrature data passed by value (kernel arguments live in the constant
// kernarg segment; every index into them is wave-uniform).
struct QuadArg {
    int nq;
    double qw[8];
    double ref[144];
};

QuadArg make_quad_arg(const mpk_quadrature* q, bool curl) {
    QuadArg a;
    std::memset(&a, 0, sizeof(a));
    a.nq = q->nq;
    for (int i = 0; i < q->nq; ++i) a.qw[i] = q->qw[i];
    const double* src = curl ? q->ref_curl : q->ref_val;
    for (int i = 0; i < 18 * q->nq; ++i) a.ref[i] = src[i];
    return a;
}

// compute_jacobian_3d, src/mex/MexElemUtils.hpp:35-49 (J row-major J[3*i+j]).
__device__ __forceinline__ double jacobian_3d(const double* p, double* J) {
#pragma clang fp contract(off)
    J[0] = p[3] - p[0]; J[3] = p[4] - p[1]; J[6] = p[5] - p[2];
    J[1] = p[6] - p[0]; J[4] = p[7] - p[1]; J[7] = p[8] - p[2];
    J[2] = p[9] - p[0]; J[5] = p[10] - p[1]; J[8] = p[11] - p[2];
    return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
           J[2] * (J[3] * J[7] - J[4] * J[6]);
}

// compute_inverse_3x3, src/mex/MexElemUtils.hpp:52-66.
__device__ __forceinline__ void inverse_3x3(const double* J, double det, double* Ji) {
#pragma clang fp contract(off)
    const double invDet = 1.0 / det;
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) * invDet;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * invDet;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * invDet;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * invDet;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * invDet;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * invDet;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * invDet;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * invDet;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * invDet;
}

// Gathers the 4 nodes of element e (T stored as doubles, 1-based).  Sets
// *bad when a node id falls outside [1, nn].
__device__ __forceinline__ void gather_nodes(const double* __restrict__ P, int64_t nn,
                                             const double* __restrict__ T, int64_t e,
                                             double* p, bool& bad) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        long node = (long)T[n + e * 4] - 1;
        if (node < 0 || node >= nn) {
            bad = true;
            node = 0;
        }
        p[3 * n + 0] = P[node * 3 + 0];
        p[3 * n + 1] = P[node * 3 + 1];
        p[3 * n + 2] = P[node * 3 + 2];
    }
}

// Writes 36 values per lane through LDS into out[36*e0 .. 36*(e0+count)).
__device__ __forceinline__ void stage_store36(double* stage, const double (&v)[36],
                                              double* __restrict__ out, int64_t e0, int count,
                                              bool active) {
    const int lane = threadIdx.x;
    if (active) {
#pragma unroll
        for (int k = 0; k < 36; ++k) stage[lane * kStageStride + k] = v[k];
    }
    __syncthreads();
    const int total = count * 36;
    for (int i = lane; i < total; i += kBlock) {
        const int el = i / 36, k = i - el * 36;
        out[e0 * 36 + i] = stage[el * kStageStride + k];
    }
    __syncthreads();
}

__device__ __forceinline__ void stage_store6(double* stage, const double (&v)[6],
                                             double* __restrict__ out, int64_t e0, int count,
                                             bool active) {
    const int lane = threadIdx.x;
    if (active) {
#pragma unroll
        for (int k = 0; k < 6; ++k) stage[lane * 7 + k] = v[k];
    }
    __syncthreads();
    const int total = count * 6;
    for (int i = lane; i < total; i += kBlock) {
        const int el = i / 6, k = i - el * 6;
        out[e0 * 6 + i] = stage[el * 7 + k];
    }
    __syncthreads();
}

// Writes N values per lane through LDS into out[N*e0 .. N*(e0+count))
// (stride N + 1: odd, against LDS bank conflicts).
template <int N>
__device__ __forceinline__ void stage_store(double* stage, const double (&v)[N],
                                            double* __restrict__ out, int64_t e0, int count,
                                            bool active) {
    const int lane = threadIdx.x;
    if (active) {
#pragma unroll
        for (int k = 0; k < N; ++k) stage[lane * (N + 1) + k] = v[k];
    }
    __syncthreads();
    const int total = count * N;
    for (int i = lane; i < total; i += kBlock) {
        const int el = i / N, k = i - el * N;
        out[e0 * N + i] = stage[el * (N + 1) + k];
    }
    __syncthreads();
}

// y = J^{-T} g for a reference row g (MexElemUtils multVecTrans), Ji row-major.
__device__ __forceinline__ void jit_row(const double* Ji, double g0, double g1, double g2,
                                        double* y) {
#pragma clang fp contract(off)
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = Ji[0 + a] * g0 + Ji[3 + a] * g1 + Ji[6 + a] * g2;
}

// Fills I, J, V for a finished block K[r][c] (reference output loop:
// assemble_curl_curl_kernel_mex.cpp:104-120).
__device__ __forceinline__ void emit_block(double* stage, const double (&K)[6][6],
                                           const double* __restrict__ D,
                                           const double* __restrict__ S, int64_t e, int64_t e0,
                                           int count, bool active, double* __restrict__ Io,
                                           double* __restrict__ Jo, double* __restrict__ Vo) {
#pragma clang fp contract(off)
    double s[6], d[6];
    if (active) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            s[k] = S[k + e * 6];
            d[k] = (double)(long)D[k + e * 6];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] = d[k] = 0.0;
    }
    double v[36];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = d[r];
    stage_store36(stage, v, Io, e0, count, active);
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = d[c];
    stage_store36(stage, v, Jo, e0, count, active);
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = K[r][c] * s[r] * s[c];
    stage_store36(stage, v, Vo, e0, count, active);
}

__global__ void __launch_bounds__(kBlock)
k_curl_curl_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T,
                int64_t ne, const double* __restrict__ D, const double* __restrict__ S,
                const double* __restrict__ Nu, QuadArg qa, double* __restrict__ Io,
                double* __restrict__ Jo, double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * kStageStride];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    double K[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) K[i][j] = 0.0;
    if (active) {
        double p[12], J[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        if (bad) atomicOr(err, 1);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        const double invDetJ = 1.0 / detJ;
        const double nu = Nu[e];
        for (int q = 0; q < qa.nq; ++q) {
            const double w = qa.qw[q] * absDetJ;
            double B[3][6];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const double r0 = qa.ref[i + 0 + 18 * q];
                const double r1 = qa.ref[i + 6 + 18 * q];
                const double r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double t = J[3 * a + 0] * r0 + J[3 * a + 1] * r1 + J[3 * a + 2] * r2;
                    B[a][i] = t * invDetJ;
                }
            }
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    const double dot = B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j];
                    K[i][j] += dot * nu * w;
                }
        }
    }
    emit_block(stage, K, D, S, e, e0, count, active, Io, Jo, Vo);
}

__global__ void __launch_bounds__(kBlock)
k_mass_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T, int64_t ne,
           const double* __restrict__ D, const double* __restrict__ S,
           const double* __restrict__ C, int has_c, QuadArg qa, double* __restrict__ Io,
           double* __restrict__ Jo, double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * kStageStride];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    double M[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) M[i][j] = 0.0;
    if (active) {
        double p[12], J[9], Ji[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        if (bad) atomicOr(err, 1);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        inverse_3x3(J, detJ, Ji);
        const double sigma = has_c ? C[e] : 1.0;
        if (fabs(sigma) > 1e-12) {
            for (int q = 0; q < qa.nq; ++q) {
                const double w = qa.qw[q] * absDetJ;
                double N[3][6];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    const double r0 = qa.ref[i + 0 + 18 * q];
                    const double r1 = qa.ref[i + 6 + 18 * q];
                    const double r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        N[a][i] = Ji[0 + a] * r0 + Ji[3 + a] * r1 + Ji[6 + a] * r2;
                }
#pragma unroll
                for (int i = 0; i < 6; ++i)
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        const double dot =
                            N[0][i] * N[0][j] + N[1][i] * N[1][j] + N[2][i] * N[2][j];
                        M[i][j] += dot * sigma * w;
                    }
            }
        }
    }
    emit_block(stage, M, D, S, e, e0, count, active, Io, Jo, Vo);
}

// B-H tables staged in LDS when they fit (kMatLds doubles), else read from HBM.
constexpr int kMatLds = 2048;

template <bool kLds>
__global__ void __launch_bounds__(kBlock)
k_jacobian_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T,
               int64_t ne, const double* __restrict__ D, const double* __restrict__ S,
               const double* __restrict__ Tags, const double* __restrict__ SolA, int64_t nsol,
               QuadArg qa, MaterialView mat, int ntags, int nbrk, int ncoe, int calcJ,
               double* __restrict__ Io, double* __restrict__ Jo, double* __restrict__ Vo,
               double* __restrict__ Ridx, double* __restrict__ Rval, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * kStageStride];
    __shared__ double lds_tab[kLds ? kMatLds : 1];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    if (kLds) {
        for (int i = threadIdx.x; i < nbrk; i += kBlock) lds_tab[i] = mat.brk[i];
        for (int i = threadIdx.x; i < ncoe; i += kBlock) lds_tab[nbrk + i] = mat.coe[i];
        __syncthreads();
    }
    double K[6][6];
    double R[6];
    double dofs[6], s[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        R[i] = 0.0;
        dofs[i] = 0.0;
        s[i] = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) K[i][j] = 0.0;
    }
    if (active) {
        double p[12], J[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        const double invDetJ = 1.0 / detJ;
        double a[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const int dof = (int)D[i + e * 6];
            dofs[i] = (double)dof;
            s[i] = S[i + e * 6];
            const int gid = dof - 1;
            if (gid >= nsol) bad = true;
            a[i] = (gid >= 0 && gid < nsol) ? SolA[gid] * s[i] : 0.0;
        }
        int tag = (int)Tags[e];
        if (tag < 1 || tag > ntags) {  // UB in the reference; an index error here
            bad = true;
            tag = 1;
        }
        if (bad) atomicOr(err, 1);
        double C[6][3];
        for (int q = 0; q < qa.nq; ++q) {
            const double w = qa.qw[q] * absDetJ;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const double r0 = qa.ref[i + 0 + 18 * q];
                const double r1 = qa.ref[i + 6 + 18 * q];
                const double r2 = qa.ref[i + 12 + 18 * q];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double t = J[3 * k + 0] * r0 + J[3 * k + 1] * r1 + J[3 * k + 2] * r2;
                    C[i][k] = t * invDetJ;
                }
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
            if (kLds)
                eval_nu(Bsq, tag, mat, lds_tab, lds_tab + nbrk, nu, dnu);
            else
                eval_nu(Bsq, tag, mat, mat.brk, mat.coe, nu, dnu);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const double BCi = B0 * C[i][0] + B1 * C[i][1] + B2 * C[i][2];
                R[i] += nu * BCi * w;
                if (calcJ) {
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        const double CiCj = C[i][0] * C[j][0] + C[i][1] * C[j][1] + C[i][2] * C[j][2];
                        const double BCj = B0 * C[j][0] + B1 * C[j][1] + B2 * C[j][2];
                        const double val = nu * CiCj + 2.0 * dnu * BCi * BCj;
                        K[i][j] += val * w;
                    }
                }
            }
        }
    }
    double rv[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) rv[i] = R[i] * s[i];
    stage_store6(stage, dofs, Ridx, e0, count, active);
    stage_store6(stage, rv, Rval, e0, count, active);
    if (calcJ) {
        double v[36];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int i = 0; i < 6; ++i) v[6 * j + i] = dofs[i];
        stage_store36(stage, v, Io, e0, count, active);
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int i = 0; i < 6; ++i) v[6 * j + i] = dofs[j];
        stage_store36(stage, v, Jo, e0, count, active);
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int i = 0; i < 6; ++i) v[6 * j + i] = K[i][j] * s[i] * s[j];
        stage_store36(stage, v, Vo, e0, count, active);
    }
}

// Reference gradient table of the P1 nodal functions, [4 x 3 x nq] at
// i + 4 c + 12 q (lagrange_tet_p1.m:22-38), passed by value.
struct GradArg {
    double g[96];
};

// K = (sigma grad phi_i, grad phi_j): slot 16 e + 4 c + r = (dofs[r], dofs[c], K[r][c]).
__global__ void __launch_bounds__(kBlock)
k_scalar_laplacian_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T,
                       int64_t ne, const double* __restrict__ D, const double* __restrict__ C,
                       QuadArg qa, GradArg ga, double* __restrict__ Io, double* __restrict__ Jo,
                       double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * 17];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    double K[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) K[i][j] = 0.0;
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
        double p[12], J[9], Ji[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        if (bad) atomicOr(err, 1);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        inverse_3x3(J, detJ, Ji);
        const double sigma = C[e];
        if (fabs(sigma) > 1e-12) {
            for (int q = 0; q < qa.nq; ++q) {
                const double w = qa.qw[q] * absDetJ;
                double G[3][4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    double y[3];
                    jit_row(Ji, ga.g[i + 0 + 12 * q], ga.g[i + 4 + 12 * q], ga.g[i + 8 + 12 * q], y);
                    G[0][i] = y[0];
                    G[1][i] = y[1];
                    G[2][i] = y[2];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double dot =
                            G[0][i] * G[0][j] + G[1][i] * G[1][j] + G[2][i] * G[2][j];
                        K[i][j] += dot * sigma * w;
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = (double)(long)D[k + e * 4];
    }
    double v[16];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[4 * c + r] = d[r];
    stage_store<16>(stage, v, Io, e0, count, active);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[4 * c + r] = d[c];
    stage_store<16>(stage, v, Jo, e0, count, active);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[4 * c + r] = K[r][c];
    stage_store<16>(stage, v, Vo, e0, count, active);
}

// C = (sigma N_i, grad phi_j), 6 edge rows x 4 node columns:
// slot 24 e + 6 c + r = (dofsE[r], dofsN[c], C[r][c] * s[r]).
__global__ void __launch_bounds__(kBlock)
k_coupling_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T,
               int64_t ne, const double* __restrict__ DE, const double* __restrict__ S,
               const double* __restrict__ DN, const double* __restrict__ C, QuadArg qa,
               GradArg ga, double* __restrict__ Io, double* __restrict__ Jo,
               double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * 25];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    double Cl[6][4];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Cl[i][j] = 0.0;
    double re[6] = {0, 0, 0, 0, 0, 0}, sg[6] = {0, 0, 0, 0, 0, 0}, cn[4] = {0, 0, 0, 0};
    if (active) {
        double p[12], J[9], Ji[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        if (bad) atomicOr(err, 1);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        inverse_3x3(J, detJ, Ji);
        const double sigma = C[e];
        if (fabs(sigma) > 1e-12) {
            for (int q = 0; q < qa.nq; ++q) {
                const double w = qa.qw[q] * absDetJ;
                double N[3][6], G[3][4];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double y[3];
                    jit_row(Ji, qa.ref[i + 0 + 18 * q], qa.ref[i + 6 + 18 * q],
                            qa.ref[i + 12 + 18 * q], y);
                    N[0][i] = y[0];
                    N[1][i] = y[1];
                    N[2][i] = y[2];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    double y[3];
                    jit_row(Ji, ga.g[i + 0 + 12 * q], ga.g[i + 4 + 12 * q], ga.g[i + 8 + 12 * q], y);
                    G[0][i] = y[0];
                    G[1][i] = y[1];
                    G[2][i] = y[2];
                }
#pragma unroll
                for (int i = 0; i < 6; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double dot =
                            N[0][i] * G[0][j] + N[1][i] * G[1][j] + N[2][i] * G[2][j];
                        Cl[i][j] += dot * sigma * w;
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            re[k] = (double)(long)DE[k + e * 6];
            sg[k] = S[k + e * 6];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) cn[k] = (double)(long)DN[k + e * 4];
    }
    double v[24];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = re[r];
    stage_store<24>(stage, v, Io, e0, count, active);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = cn[c];
    stage_store<24>(stage, v, Jo, e0, count, active);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) v[6 * c + r] = Cl[r][c] * sg[r];
    stage_store<24>(stage, v, Vo, e0, count, active);
}

// Winding source F_i = (Scale J_dir, N_i) on the elements of one coil
// (tag == coil_id as ints); other elements give (0, 0) pairs.
__global__ void __launch_bounds__(kBlock)
k_winding_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T,
              int64_t ne, const double* __restrict__ D, const double* __restrict__ S,
              const double* __restrict__ Tags, int coil, const double* __restrict__ Dir,
              double u0, double u1, double u2, double scale, QuadArg qa,
              double* __restrict__ Io, double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double stage[kBlock * 7];
    const int64_t e0 = (int64_t)blockIdx.x * kBlock;
    const int64_t e = e0 + threadIdx.x;
    const int count = (int)((ne - e0) < kBlock ? (ne - e0) : kBlock);
    const bool active = e < ne;
    double idx[6] = {0, 0, 0, 0, 0, 0}, val[6] = {0, 0, 0, 0, 0, 0};
    if (active && (int)Tags[e] == coil) {
        double p[12], J[9], Ji[9];
        bool bad = false;
        gather_nodes(P, nn, T, e, p, bad);
        if (bad) atomicOr(err, 1);
        const double detJ = jacobian_3d(p, J);
        const double absDetJ = fabs(detJ);
        inverse_3x3(J, detJ, Ji);
        double jp[3];
        if (Dir) {
            jp[0] = Dir[0 + e * 3];
            jp[1] = Dir[1 + e * 3];
            jp[2] = Dir[2 + e * 3];
        } else {
            jp[0] = u0;
            jp[1] = u1;
            jp[2] = u2;
        }
        jp[0] *= scale;
        jp[1] *= scale;
        jp[2] *= scale;
        double jt[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            jt[i] = Ji[3 * i + 0] * jp[0] + Ji[3 * i + 1] * jp[1] + Ji[3 * i + 2] * jp[2];
        double F[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < qa.nq; ++q) {
            const double w = qa.qw[q] * absDetJ;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                const double v = jt[0] * qa.ref[i + 0 + 18 * q] + jt[1] * qa.ref[i + 6 + 18 * q] +
                                 jt[2] * qa.ref[i + 12 + 18 * q];
                F[i] += v * w;
            }
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            idx[i] = (double)(long)D[i + e * 6];
            val[i] = F[i] * S[i + e * 6];
        }
    }
    stage_store<6>(stage, idx, Io, e0, count, active);
    stage_store<6>(stage, val, Vo, e0, count, active);
}

constexpr int kHbMaxHarm = 64;  // 24 * 64 doubles per lane: one lane per 12 KB block

// Harmonic-balance residual and DC-reluctivity block Jacobian, one element
// per lane.  The lane's phasors A (real, imaginary) and residual accumulators
// (real, imaginary), 6 x H each, live in LDS at lane * 24 H; A_time of a
// sample is formed where it is used (the same sums, in the same order, as the
// reference's precomputed table).
struct HbArgs {
    const double* sol_r;
    const double* sol_i;  // may be NULL
    int64_t num_dofs;
    int H;
    const double* dmat_r;  // [Nt x H]
    const double* dmat_i;  // may be NULL
    int Nt;
    const double* pmat_r;  // [H x Nt]
    const double* pmat_i;  // may be NULL
    const double* scalings;
    int request_j;
};

__global__ void __launch_bounds__(kBlock)
k_hbfem_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T, int64_t ne,
            const double* __restrict__ D, const double* __restrict__ S,
            const double* __restrict__ Tags, HbArgs hb, QuadArg qa, MaterialView mat, int ntags,
            double* __restrict__ Io, double* __restrict__ Jo, double* __restrict__ Vo,
            double* __restrict__ Rrows, double* __restrict__ Rcols, double* __restrict__ RvR,
            double* __restrict__ RvI, int* __restrict__ err) {
#pragma clang fp contract(off)
    extern __shared__ double hb_lds[];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int H = hb.H;
    double* AR = hb_lds + (size_t)threadIdx.x * 24 * H;
    double* AI = AR + 6 * H;
    double* ReR = AR + 12 * H;
    double* ReI = AR + 18 * H;
    bool bad = false;
    double dofs[6], s[6];
    for (int k = 0; k < 6; ++k) {
        dofs[k] = D[k + e * 6];
        s[k] = S[k + e * 6];
        const int gid = (int)dofs[k] - 1;
        if (gid >= hb.num_dofs) bad = true;
        for (int h = 0; h < H; ++h) {
            const int64_t idx = (int64_t)gid + (int64_t)h * hb.num_dofs;
            const bool ok = idx >= 0 && idx < hb.num_dofs * H;
            AR[k + h * 6] = (ok ? hb.sol_r[idx] : 0.0) * s[k];
            AI[k + h * 6] = hb.sol_i ? (ok ? hb.sol_i[idx] : 0.0) * s[k] : 0.0;
            ReR[k + h * 6] = 0.0;
            ReI[k + h * 6] = 0.0;
        }
    }
    int tag = (int)Tags[e];
    if (tag < 1 || tag > ntags) {  // UB in the reference; an index error here
        bad = true;
        tag = 1;
    }
    double p[12], J[9];
    gather_nodes(P, nn, T, e, p, bad);
    if (bad) atomicOr(err, 1);
    const double detJ = jacobian_3d(p, J);
    const double invDetJ = 1.0 / detJ;
    const double absDetJ = fabs(detJ);
    const bool imag_a = hb.sol_i && hb.dmat_i;
    double Ke[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) Ke[i] = 0.0;
    for (int q = 0; q < qa.nq; ++q) {
        const double w = qa.qw[q] * absDetJ;
        double C[18];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double cx = qa.ref[i + 0 + 18 * q], cy = qa.ref[i + 6 + 18 * q],
                         cz = qa.ref[i + 12 + 18 * q];
            C[i + 0] = (cx * J[0] + cy * J[1] + cz * J[2]) * invDetJ;
            C[i + 6] = (cx * J[3] + cy * J[4] + cz * J[5]) * invDetJ;
            C[i + 12] = (cx * J[6] + cy * J[7] + cz * J[8]) * invDetJ;
        }
        double sum_nu = 0.0;
        for (int t = 0; t < hb.Nt; ++t) {
            double Bx = 0.0, By = 0.0, Bz = 0.0;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double v = 0.0;
                for (int h = 0; h < H; ++h) {
                    double term = AR[i + h * 6] * hb.dmat_r[t + (int64_t)h * hb.Nt];
                    if (imag_a) term -= AI[i + h * 6] * hb.dmat_i[t + (int64_t)h * hb.Nt];
                    v += term;
                }
                Bx += C[i + 0] * v;
                By += C[i + 6] * v;
                Bz += C[i + 12] * v;
            }
            const double bsq = Bx * Bx + By * By + Bz * Bz;
            double nu, dnu;
            eval_nu(bsq, tag, mat, mat.brk, mat.coe, nu, dnu);
            sum_nu += nu;
            const double Hx = nu * Bx, Hy = nu * By, Hz = nu * Bz;
            for (int h = 0; h < H; ++h) {
                const double pr = hb.pmat_r[h + (int64_t)t * H];
                const double pi = hb.pmat_i ? hb.pmat_i[h + (int64_t)t * H] : 0.0;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    const double dot = C[i + 0] * Hx + C[i + 6] * Hy + C[i + 12] * Hz;
                    ReR[i + h * 6] += w * dot * pr;
                    if (hb.pmat_i) ReI[i + h * 6] += w * dot * pi;
                }
            }
        }
        if (hb.request_j) {
            const double nu_dc = sum_nu / (double)hb.Nt;
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    const double dot =
                        C[r + 0] * C[c + 0] + C[r + 6] * C[c + 6] + C[r + 12] * C[c + 12];
                    Ke[r + c * 6] += w * nu_dc * dot;
                }
        }
    }
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int64_t o = (int64_t)6 * H * e + 6 * h + k;
            Rrows[o] = dofs[k];
            Rcols[o] = (double)(h + 1);
            RvR[o] = ReR[k + h * 6] * s[k];
            RvI[o] = ReI[k + h * 6] * s[k];
        }
    if (hb.request_j) {
        for (int h = 0; h < H; ++h) {
            const double f = hb.scalings[h];
            const double off = (double)((int64_t)h * hb.num_dofs);
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    const int64_t o = (int64_t)36 * H * e + 36 * h + 6 * c + r;
                    const double val = Ke[r + c * 6] * f;
                    Io[o] = dofs[r] + off;
                    Jo[o] = dofs[c] + off;
                    Vo[o] = val * s[r] * s[c];
                }
        }
    }
}

int check_quad(const mpk_quadrature* q) {
    MPK_REQUIRE(q != nullptr, MPK_ERR_ARGS, "quadrature table is NULL");
    MPK_REQUIRE(q->nq >= 1 && q->nq <= 8, MPK_ERR_ARGS, "nq=%d outside [1, 8]", q->nq);
    return MPK_OK;
}

// Runs `launch(err_dev)` and reports an index error raised by the kernel.
template <class F>
int run_checked(hipStream_t st, F&& launch) {
    DevBuf err;
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    launch(err.as<int>());
    MPK_HIP_TRY(hipGetLastError());
    int h = 0;
    MPK_HIP_TRY(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    MPK_REQUIRE(h == 0, MPK_ERR_INDEX, "node or DOF index outside its range");
    return MPK_OK;
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_abi_version(void) { return 1; }

const char* mpk_last_error(void) { return mpk::g_err.c_str(); }

int mpk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mpk_curl_curl_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                      const double* cell_dofs, const double* signs, const double* nu,
                      const mpk_quadrature* quad, double* I_out, double* J_out,
                      double* V_out, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && nu && I_out && J_out && V_out, MPK_ERR_ARGS,
                "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    const QuadArg qa = make_quad_arg(quad, true);
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_curl_curl_coo, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, st, P,
                           nn, T, ne, cell_dofs, signs, nu, qa, I_out, J_out, V_out, err);
    });
}

int mpk_mass_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                 const double* cell_dofs, const double* signs, const double* coeffs,
                 int64_t num_coeffs, const mpk_quadrature* quad, double* I_out,
                 double* J_out, double* V_out, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && I_out && J_out && V_out, MPK_ERR_ARGS,
                "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    const int has_c = (coeffs != nullptr && num_coeffs == ne) ? 1 : 0;
    const QuadArg qa = make_quad_arg(quad, false);
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_mass_coo, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, st, P, nn,
                           T, ne, cell_dofs, signs, coeffs, has_c, qa, I_out, J_out, V_out, err);
    });
}

int mpk_jacobian_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                     const double* cell_dofs, const double* signs, const double* tags,
                     const double* sol_a, int64_t num_sol, const mpk_quadrature* quad,
                     const mpk_material* mat, int calc_j, double* I_out, double* J_out,
                     double* V_out, double* R_idx, double* R_val, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(mat != nullptr, MPK_ERR_ARGS, "material table is NULL");
    MPK_REQUIRE(ne >= 0 && nn >= 0 && num_sol >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && tags && R_idx && R_val, MPK_ERR_ARGS,
                "NULL array argument");
    MPK_REQUIRE(!calc_j || (I_out && J_out && V_out), MPK_ERR_ARGS, "NULL J output");
    MPK_REQUIRE(num_sol == 0 || sol_a, MPK_ERR_ARGS, "NULL SolA");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    MPK_REQUIRE(mat->num_tags >= 1 && mat->linear_nu && mat->is_nonlinear && mat->max_bsq &&
                    mat->spline_start && mat->spline_count && mat->coef_start,
                MPK_ERR_ARGS, "material table without tags");
    MaterialView mv{mat->linear_nu, mat->is_nonlinear, mat->max_bsq,   mat->breaks,
                    mat->coefs,     mat->spline_start, mat->spline_count, mat->coef_start};
    const int nbrk = (int)mat->num_breaks, ncoe = (int)mat->num_coefs;
    const int ntags = (int)mat->num_tags;
    const QuadArg qa = make_quad_arg(quad, true);
    hipStream_t st = as_stream(stream);
    const bool lds = (int64_t)nbrk + ncoe <= kMatLds;
    return run_checked(st, [&](int* err) {
        if (lds)
            hipLaunchKernelGGL(k_jacobian_coo<true>, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0,
                               st, P, nn, T, ne, cell_dofs, signs, tags, sol_a, num_sol, qa, mv,
                               ntags, nbrk, ncoe, calc_j, I_out, J_out, V_out, R_idx, R_val, err);
        else
            hipLaunchKernelGGL(k_jacobian_coo<false>, dim3(grid_for(ne, kBlock)), dim3(kBlock),
                               0, st, P, nn, T, ne, cell_dofs, signs, tags, sol_a, num_sol, qa,
                               mv, ntags, nbrk, ncoe, calc_j, I_out, J_out, V_out, R_idx, R_val, err);
    });
}

int mpk_scalar_laplacian_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                             const double* node_dofs, const double* coeffs,
                             const mpk_quadrature* quad, const double* grad_ref,
                             double* I_out, double* J_out, double* V_out, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && node_dofs && coeffs && grad_ref && I_out && J_out && V_out,
                MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    const QuadArg qa = make_quad_arg(quad, false);
    GradArg ga;
    std::memset(&ga, 0, sizeof(ga));
    for (int i = 0; i < 12 * quad->nq; ++i) ga.g[i] = grad_ref[i];
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_scalar_laplacian_coo, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, st,
                           P, nn, T, ne, node_dofs, coeffs, qa, ga, I_out, J_out, V_out, err);
    });
}

int mpk_coupling_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                     const double* edge_dofs, const double* signs, const double* node_dofs,
                     const double* coeffs, const mpk_quadrature* quad, const double* grad_ref,
                     double* I_out, double* J_out, double* V_out, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && edge_dofs && signs && node_dofs && coeffs && grad_ref && I_out &&
                    J_out && V_out,
                MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    const QuadArg qa = make_quad_arg(quad, false);
    GradArg ga;
    std::memset(&ga, 0, sizeof(ga));
    for (int i = 0; i < 12 * quad->nq; ++i) ga.g[i] = grad_ref[i];
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_coupling_coo, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, st, P, nn,
                           T, ne, edge_dofs, signs, node_dofs, coeffs, qa, ga, I_out, J_out, V_out,
                           err);
    });
}

int mpk_winding_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                    const double* cell_dofs, const double* signs, const double* tags,
                    double coil_id, const double* dir_field, const double* uniform_dir,
                    double scale, const mpk_quadrature* quad, double* I_out, double* V_out,
                    void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && tags && uniform_dir && I_out && V_out,
                MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    const QuadArg qa = make_quad_arg(quad, false);
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_winding_coo, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, st, P, nn,
                           T, ne, cell_dofs, signs, tags, (int)coil_id, dir_field, uniform_dir[0],
                           uniform_dir[1], uniform_dir[2], scale, qa, I_out, V_out, err);
    });
}

int mpk_hbfem_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                  const double* cell_dofs, const double* signs, const double* tags,
                  const double* sol_r, const double* sol_i, int64_t num_dofs, int32_t num_harm,
                  const double* dmat_r, const double* dmat_i, int32_t num_time,
                  const double* pmat_r, const double* pmat_i, const mpk_quadrature* quad,
                  const mpk_material* mat, int request_j, const double* scalings,
                  double* I_out, double* J_out, double* V_out, double* R_rows, double* R_cols,
                  double* R_val_r, double* R_val_i, void* stream) {
    clear_error();
    if (int rc = check_quad(quad)) return rc;
    MPK_REQUIRE(mat != nullptr, MPK_ERR_ARGS, "material table is NULL");
    MPK_REQUIRE(ne >= 0 && nn >= 0 && num_dofs >= 0, MPK_ERR_ARGS, "negative size");
    MPK_REQUIRE(num_harm >= 1 && num_harm <= kHbMaxHarm, MPK_ERR_LIMIT,
                "%d harmonics outside [1, %d]", num_harm, kHbMaxHarm);
    MPK_REQUIRE(num_time >= 1, MPK_ERR_ARGS, "no time samples");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && tags && sol_r && dmat_r && pmat_r && R_rows &&
                    R_cols && R_val_r && R_val_i,
                MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(!request_j || (I_out && J_out && V_out && scalings), MPK_ERR_ARGS,
                "NULL Jacobian output or scalings");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    MPK_REQUIRE(mat->num_tags >= 1 && mat->linear_nu && mat->is_nonlinear && mat->max_bsq &&
                    mat->spline_start && mat->spline_count && mat->coef_start,
                MPK_ERR_ARGS, "material table without tags");
    MaterialView mv{mat->linear_nu, mat->is_nonlinear, mat->max_bsq,   mat->breaks,
                    mat->coefs,     mat->spline_start, mat->spline_count, mat->coef_start};
    HbArgs hb{sol_r, sol_i, num_dofs, num_harm, dmat_r, dmat_i, num_time, pmat_r, pmat_i,
              scalings, request_j ? 1 : 0};
    const QuadArg qa = make_quad_arg(quad, true);
    // lanes per block so that 24 H doubles per lane fit in 64 KB of LDS
    int lanes = kBlock;
    while (lanes > 1 && (size_t)lanes * 24 * num_harm * sizeof(double) > 65536) lanes /= 2;
    const size_t lds = (size_t)lanes * 24 * num_harm * sizeof(double);
    hipStream_t st = as_stream(stream);
    return run_checked(st, [&](int* err) {
        hipLaunchKernelGGL(k_hbfem_coo, dim3(grid_for(ne, lanes)), dim3(lanes), lds, st, P, nn, T,
                           ne, cell_dofs, signs, tags, hb, qa, mv, (int)mat->num_tags, I_out,
                           J_out, V_out, R_rows, R_cols, R_val_r, R_val_i, err);
    });
}

}  // extern "C"

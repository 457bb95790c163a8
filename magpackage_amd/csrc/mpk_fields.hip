// Per-element field kernels around the assembly path, FP64 in the
// reference's operation order (no FMA contraction):
//
//   k_source_coo   assemble_source_kernel_mex (src/mex/assemble_source_kernel_mex.cpp:16-120):
//                  6 (dof, F_i * s_i) pairs per element, slot 6e + i.
//   k_csc_source   the same values summed per DOF, i.e. sparse(I(valid), 1, V(valid), N, 1)
//                  of src/Assembler.m:87-89: one lane per DOF walks the plan's incidence
//                  list (ascending element order == triplet order), no atomics.
//   k_element_b    PostProcessor.computeElementB (src/post/PostProcessor.m:64-126):
//                  B_e = sum_i (C_ref J' / det)(i, :) * A(dof_i) s_i.
#include <cmath>
#include <cstring>

#include "mpk_csc_exact.hpp"
#include "mpk_plan.hpp"

namespace mpk {
namespace {

using namespace csc;

constexpr int kB = 256;

struct ValArg {
    int nq;
    double qw[8];
    double ref[144];
};

// J^{-1} J_s of one element (assemble_source_kernel_mex.cpp:57-76): geometry
// in the MEX's order, returns |det| through absdet.
__device__ __forceinline__ void source_setup(const double* p, const double* js, double* jt,
                                             double& absdet) {
#pragma clang fp contract(off)
    double J[9], Ji[9];
    J[0] = p[3] - p[0]; J[3] = p[4] - p[1]; J[6] = p[5] - p[2];
    J[1] = p[6] - p[0]; J[4] = p[7] - p[1]; J[7] = p[8] - p[2];
    J[2] = p[9] - p[0]; J[5] = p[10] - p[1]; J[8] = p[11] - p[2];
    const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                       J[2] * (J[3] * J[7] - J[4] * J[6]);
    absdet = fabs(det);
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
#pragma unroll
    for (int i = 0; i < 3; ++i) jt[i] = Ji[3 * i + 0] * js[0] + Ji[3 * i + 1] * js[1] + Ji[3 * i + 2] * js[2];
}

// F_i of one element for local edge i (before the sign), :80-101.
__device__ __forceinline__ double source_entry(const double* jt, const double* js, double absdet,
                                               int i, const ValArg& qa) {
#pragma clang fp contract(off)
    double f = 0.0;
    if (fabs(js[0]) > 1e-12 || fabs(js[1]) > 1e-12 || fabs(js[2]) > 1e-12) {
        for (int q = 0; q < qa.nq; ++q) {
            const double w = qa.qw[q] * absdet;
            const int qo = 18 * q;
            const double v = jt[0] * qa.ref[i + 0 + qo] + jt[1] * qa.ref[i + 6 + qo] +
                             jt[2] * qa.ref[i + 12 + qo];
            f += v * w;
        }
    }
    return f;
}

__global__ void __launch_bounds__(kB)
k_source_coo(const double* __restrict__ P, int64_t nn, const double* __restrict__ T, int64_t ne,
             const double* __restrict__ D, const double* __restrict__ S,
             const double* __restrict__ Js, ValArg qa, double* __restrict__ Io,
             double* __restrict__ Vo, int* __restrict__ err) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne) return;
    double p[12];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        long node = (long)T[n + e * 4] - 1;
        if (node < 0 || node >= nn) {
            atomicOr(err, 1);
            node = 0;
        }
        p[3 * n + 0] = P[node * 3 + 0];
        p[3 * n + 1] = P[node * 3 + 1];
        p[3 * n + 2] = P[node * 3 + 2];
    }
    const double js[3] = {Js[3 * e + 0], Js[3 * e + 1], Js[3 * e + 2]};
    double jt[3], absdet;
    source_setup(p, js, jt, absdet);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double f = source_entry(jt, js, absdet, i, qa);
        Io[6 * e + i] = (double)(long)D[i + 6 * e];
        Vo[6 * e + i] = f * S[i + 6 * e];
    }
}

// One lane per DOF (plan column); incidences in ascending element order.
__global__ void __launch_bounds__(kB)
k_csc_source(const int32_t* __restrict__ inc_ptr, const uint32_t* __restrict__ inc_slot,
             const int4* __restrict__ T, const double* __restrict__ P,
             const double* __restrict__ Js, int64_t n, ValArg qa, double* __restrict__ F) {
#pragma clang fp contract(off)
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    double acc = 0.0;
    for (int k = inc_ptr[j]; k < inc_ptr[j + 1]; ++k) {
        const uint32_t s = inc_slot[k];
        const uint32_t e = s / 6u;
        const int c = (int)(s - 6u * e);
        const int4 t = T[e];
        const NodesE nd = fetch_orig(P, t);
        const double js[3] = {Js[3 * (int64_t)e + 0], Js[3 * (int64_t)e + 1],
                              Js[3 * (int64_t)e + 2]};
        double jt[3], absdet;
        source_setup(nd.p, js, jt, absdet);
        double sg[6];
        orig_signs(t, sg);
        const double f = source_entry(jt, js, absdet, c, qa);
        acc += f * sel6(c, sg[0], sg[1], sg[2], sg[3], sg[4], sg[5]);
    }
    F[j] = acc;
}

__global__ void __launch_bounds__(kB)
k_element_b(const double* __restrict__ P, int64_t nn, const int32_t* __restrict__ T,
            const int32_t* __restrict__ D, int64_t ne, const double* __restrict__ sol,
            int64_t nsol, ValArg qc, double* __restrict__ B, int* __restrict__ err) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne) return;
    int nd[4];
    double p[12];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        int node = T[4 * e + n] - 1;
        if (node < 0 || node >= nn) {
            atomicOr(err, 1);
            node = 0;
        }
        nd[n] = node;
        p[3 * n + 0] = P[3 * (int64_t)node + 0];
        p[3 * n + 1] = P[3 * (int64_t)node + 1];
        p[3 * n + 2] = P[3 * (int64_t)node + 2];
    }
    double J[9];
    J[0] = p[3] - p[0]; J[3] = p[4] - p[1]; J[6] = p[5] - p[2];
    J[1] = p[6] - p[0]; J[4] = p[7] - p[1]; J[7] = p[8] - p[2];
    J[2] = p[9] - p[0]; J[5] = p[10] - p[1]; J[8] = p[11] - p[2];
    const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                       J[2] * (J[3] * J[7] - J[4] * J[6]);
    const double invDet = 1.0 / det;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int d = D[6 * e + i];
        if (d < 0 || d > nsol) atomicOr(err, 2);
        const double s = nd[edge_a(i)] <= nd[edge_b(i)] ? 1.0 : -1.0;
        const double a = (d > 0 && d <= nsol) ? sol[d - 1] * s : 0.0;
        const double c0 = qc.ref[i + 0], c1 = qc.ref[i + 6], c2 = qc.ref[i + 12];
        b0 += ((c0 * J[0] + c1 * J[1] + c2 * J[2]) * invDet) * a;
        b1 += ((c0 * J[3] + c1 * J[4] + c2 * J[5]) * invDet) * a;
        b2 += ((c0 * J[6] + c1 * J[7] + c2 * J[8]) * invDet) * a;
    }
    B[3 * e + 0] = b0;
    B[3 * e + 1] = b1;
    B[3 * e + 2] = b2;
}

// (nu, dnu/ds) = evaluate_nu_derivative(s, tag) (MexMaterialUtils.hpp:24-87)
// for a batch of points, tables read through the cache.
__global__ void __launch_bounds__(kB)
k_eval_nu(const double* __restrict__ s, const int32_t* __restrict__ tags, int64_t n,
          MaterialView mat, int ntags, double* __restrict__ nu, double* __restrict__ dnu,
          int* __restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    int tag = tags[i];
    if (tag < 1 || tag > ntags) {
        atomicOr(err, 1);
        tag = 1;
    }
    double a, b;
    eval_nu(s[i], tag, mat, mat.brk, mat.coe, a, b);
    nu[i] = a;
    if (dnu) dnu[i] = b;
}

ValArg make_arg(const mpk_quadrature* q, bool curl) {
    ValArg a;
    std::memset(&a, 0, sizeof(a));
    a.nq = q->nq;
    for (int i = 0; i < q->nq; ++i) a.qw[i] = q->qw[i];
    const double* src = curl ? q->ref_curl : q->ref_val;
    for (int i = 0; i < 18 * q->nq; ++i) a.ref[i] = src[i];
    return a;
}

int check_err(hipStream_t st, DevBuf& err, const char* what) {
    MPK_HIP_TRY(hipGetLastError());
    int h = 0;
    MPK_HIP_TRY(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    MPK_REQUIRE(h == 0, MPK_ERR_INDEX, "%s", what);
    return MPK_OK;
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_source_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                   const double* cell_dofs, const double* signs, const double* Js,
                   const mpk_quadrature* quad, double* I_out, double* V_out, void* stream) {
    clear_error();
    MPK_REQUIRE(quad && quad->nq >= 1 && quad->nq <= 8, MPK_ERR_ARGS, "bad quadrature table");
    MPK_REQUIRE(ne >= 0 && nn >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && signs && Js && I_out && V_out, MPK_ERR_ARGS,
                "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    hipStream_t st = as_stream(stream);
    DevBuf err;
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_source_coo, dim3(grid_for(ne, kB)), dim3(kB), 0, st, P, nn, T, ne,
                       cell_dofs, signs, Js, make_arg(quad, false), I_out, V_out,
                       err.as<int>());
    return check_err(st, err, "node index outside its range");
}

int mpk_csc_source(const mpk_plan* plan, const double* P, const double* Js,
                   const mpk_quadrature* quad, double* F, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(quad && quad->nq >= 1 && quad->nq <= 8, MPK_ERR_ARGS, "bad quadrature table");
    if (plan->n == 0) return MPK_OK;
    MPK_REQUIRE(P && Js && F, MPK_ERR_ARGS, "NULL array argument");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_csc_source, dim3(grid_for(plan->n, kB)), dim3(kB), 0, st,
                       plan->inc_ptr, plan->inc_slot, plan->T, P, Js, plan->n,
                       make_arg(quad, false), F);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_element_b(const double* P, int64_t nn, const int32_t* T, const int32_t* cell_dofs,
                  int64_t ne, const double* sol, int64_t num_sol, const mpk_quadrature* quad,
                  double* B_out, void* stream) {
    clear_error();
    MPK_REQUIRE(quad && quad->nq >= 1, MPK_ERR_ARGS, "bad quadrature table");
    MPK_REQUIRE(ne >= 0 && nn >= 0 && num_sol >= 0, MPK_ERR_ARGS, "negative size");
    if (ne == 0) return MPK_OK;
    MPK_REQUIRE(P && T && cell_dofs && sol && B_out, MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(nn >= 1, MPK_ERR_INDEX, "elements given but no nodes");
    hipStream_t st = as_stream(stream);
    DevBuf err;
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_element_b, dim3(grid_for(ne, kB)), dim3(kB), 0, st, P, nn, T,
                       cell_dofs, ne, sol, num_sol, make_arg(quad, true), B_out,
                       err.as<int>());
    return check_err(st, err, "node or DOF index outside its range");
}

int mpk_eval_nu(const double* s, const int32_t* tags, int64_t n, const mpk_material* mat,
                double* nu_out, double* dnu_out, void* stream) {
    clear_error();
    MPK_REQUIRE(mat && n >= 0, MPK_ERR_ARGS, "bad argument");
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(s && tags && nu_out && mat->num_tags >= 1, MPK_ERR_ARGS, "NULL argument");
    MaterialView mv{mat->linear_nu, mat->is_nonlinear, mat->max_bsq,    mat->breaks,
                    mat->coefs,     mat->spline_start, mat->spline_count, mat->coef_start};
    hipStream_t st = as_stream(stream);
    DevBuf err;
    MPK_HIP_TRY(err.alloc(sizeof(int)));
    MPK_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_eval_nu, dim3(grid_for(n, kB)), dim3(kB), 0, st, s, tags, n, mv,
                       (int)mat->num_tags, nu_out, dnu_out, err.as<int>());
    return check_err(st, err, "region tag outside the material table");
}

}  // extern "C"

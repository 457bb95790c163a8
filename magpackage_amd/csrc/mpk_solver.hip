// Sparse linear algebra for the transient Newton driver on gfx950.
//
// The reference solves each Newton step with MUMPS (src/solver/LinearSolver.m:22-46),
// whose binaries are absent; both arms of the learned/plain comparison use this
// solver instead.  The system is the Jacobian after BoundaryCondition.applyDirichlet
// (src/BoundaryCondition.m:81-112): rows and columns of fixed DOFs removed and a unit
// diagonal put back.  That operator is applied matrix-free from the assembled
// (symmetric, structural-pattern) CSC values and a fixed-DOF mask, so the pattern is
// never rebuilt.
//
// Preconditioned conjugate gradients with every scalar kept on the device: one
// Jacobi-PCG iteration is 5 kernels (SpMV + p.q, alpha, update + r.z and r.r,
// beta, direction), the auxiliary-space one adds G^T r and the preconditioned
// residual; no host round trip, the host checks convergence every
// `check_every` iterations.  Dot products reduce per block into a fixed slot
// array and the final sum runs in slot order in one block, so results are
// deterministic.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "mpk_common.hpp"

namespace mpk {
namespace {

constexpr int kB = 256;
constexpr int kRowLanes = 8;             // lanes cooperating on one CSR row
constexpr int kMaxPartials = 4096;       // reduction slots (grid-stride kernels)

// y = A_bc x where A is symmetric and stored CSC (== CSR of A), for a vector x
// that vanishes on the fixed DOFs (every CG search direction here does: r, z
// and p are zero on fixed rows).  Fixed rows are replaced by the identity;
// fixed columns then contribute A_ij * 0 = 0 and need no test.
//
// Segmented form: each wave owns kRows consecutive rows, streams their
// entries [colptr[r0], colptr[r0 + kRows]) with coalesced loads (64
// consecutive entries per load, a window's loads and gathers in flight),
// writes the products v_k x_j to a per-wave LDS window, and each lane then
// sums its own row's products in entry order (deterministic, independent of
// the row-length distribution).  Rows longer than the window are summed
// window by window in the same order.
// kU loads of 64 entries per wave are issued before their gathers; the LDS
// window holds kPerRow products per row (the rows of a Kuhn-mesh edge space
// hold about 16), a group of kRows rows one or two windows.
template <int kRows, int kPerRow>
struct SegWin {
    static constexpr int value = kRows * kPerRow;  // LDS products per wave
};
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row sums s_i = sum_k vals[k] x[rowind[k]] over the rows of this wave's
// groups of kRows rows (grid-stride over the groups), each handed to
// epi(i, s_i) by the lane that owns row i.  pw: this wave's LDS window.
template <int kRows, int kU, int kPerRow, class Epi>
__device__ __forceinline__ void seg_rows(const int32_t* __restrict__ colptr,
                                         const int32_t* __restrict__ rowind,
                                         const double* __restrict__ vals, int64_t n,
                                         const double* __restrict__ x, double* pw, Epi&& epi) {
    constexpr int kWin = SegWin<kRows, kPerRow>::value;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / 64) * kRows;
    for (int64_t r0 = ((int64_t)blockIdx.x * (kB / 64) + wv) * kRows; r0 < n;
         r0 += rows_per_grid) {
        const int64_t rend = r0 + kRows < n ? r0 + kRows : n;
        const bool own = lane < rend - r0;  // lane owns row r0 + lane
        const int64_t i = own ? r0 + lane : r0;
        const int rs = colptr[i];
        const int re = own ? colptr[i + 1] : rs;
        const int base = colptr[r0], top = colptr[rend];
        double s = 0.0;
        for (int w0 = base; w0 < top; w0 += kWin) {
            const int w1 = top - w0 < kWin ? top : w0 + kWin;
            for (int c = w0; c < w1; c += 64 * kU) {
                int kk[kU], jj[kU];
                double v[kU], xv[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    kk[u] = c + u * 64 + lane;
                    const int kc = kk[u] < w1 ? kk[u] : w1 - 1;  // a valid entry
                    jj[u] = rowind[kc];
                    v[u] = vals[kc];
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) xv[u] = x[jj[u]];
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (kk[u] < w1) pw[kk[u] - w0] = v[u] * xv[u];
            }
            wave_lds_sync();
            const int a = rs > w0 ? rs : w0, b = re < w1 ? re : w1;
            for (int k = a; k < b; ++k) s += pw[k - w0];
            wave_lds_sync();  // every lane has read the window before it is reused
        }
        if (own) epi(i, s);
    }
}

template <int kRows, int kU, int kPerRow>
__global__ void __launch_bounds__(kB)
k_spmv_dot(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
           const double* __restrict__ vals, const uint8_t* __restrict__ fixed, int64_t n,
           const double* __restrict__ x, double* __restrict__ y, double* __restrict__ partial) {
    __shared__ double prod[kB / 64][SegWin<kRows, kPerRow>::value];
    __shared__ double red[kB / 64];
    double acc_dot = 0.0;
    seg_rows<kRows, kU, kPerRow>(colptr, rowind, vals, n, x, prod[threadIdx.x >> 6],
                                 [&](int64_t i, double s) {
        if (fixed && fixed[i]) s = x[i];
        y[i] = s;
        acc_dot += s * x[i];
    });
    // block reduction of p.q
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc_dot += __shfl_xor(acc_dot, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc_dot;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kB / 64; ++w) t += red[w];
        partial[blockIdx.x] = t;
    }
}

// Sum of nparts partials in slot order (one block of 256 lanes, fixed tree):
// lane l adds slots l, l + 256, ... in that order (four loads issued ahead of
// their additions), then a fixed tree over the lanes.
__device__ __forceinline__ double lane_sum_fixed(const double* __restrict__ partial, int nparts) {
    double t = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * kB < nparts; i += 4 * kB) {
        const double a = partial[i], b = partial[i + kB], c = partial[i + 2 * kB],
                     d = partial[i + 3 * kB];
        t += a;
        t += b;
        t += c;
        t += d;
    }
    for (; i < nparts; i += kB) t += partial[i];
    return t;
}
__device__ double block_sum_fixed(const double* __restrict__ partial, int nparts) {
    __shared__ double sh[kB];
    sh[threadIdx.x] = lane_sum_fixed(partial, nparts);
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
// Two such sums at once (the same order for each as block_sum_fixed).
__device__ void block_sum2_fixed(const double* __restrict__ pa, const double* __restrict__ pb,
                                 int nparts, double* ra, double* rb) {
    __shared__ double sh[2][kB];
    sh[0][threadIdx.x] = lane_sum_fixed(pa, nparts);
    sh[1][threadIdx.x] = lane_sum_fixed(pb, nparts);
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    *ra = sh[0][0];
    *rb = sh[1][0];
    __syncthreads();
}

// scal[0] = rz, scal[1] = alpha, scal[2] = beta, scal[3] = |r|^2
__global__ void __launch_bounds__(kB) k_alpha(const double* __restrict__ partial, int nparts,
                                              double* __restrict__ scal) {
    const double pq = block_sum_fixed(partial, nparts);
    if (threadIdx.x == 0) scal[1] = (pq != 0.0) ? scal[0] / pq : 0.0;
}

__global__ void __launch_bounds__(kB)
k_update(int64_t n, const double* __restrict__ scal, const double* __restrict__ p,
         const double* __restrict__ q, const double* __restrict__ dinv, double* __restrict__ x,
         double* __restrict__ r, double* __restrict__ z, double* __restrict__ partial_rz,
         double* __restrict__ partial_rr) {
    __shared__ double red[2][kB / 64];
    const double alpha = scal[1];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double zi = ri * dinv[i];
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        rz += __shfl_xor(rz, off, 64);
        rr += __shfl_xor(rr, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = rz;
        red[1][threadIdx.x >> 6] = rr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kB / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        partial_rz[blockIdx.x] = a;
        partial_rr[blockIdx.x] = b;
    }
}

__global__ void __launch_bounds__(kB) k_beta(const double* __restrict__ prz,
                                             const double* __restrict__ prr, int nparts,
                                             double* __restrict__ scal) {
    double rz_new, rr;
    block_sum2_fixed(prz, prr, nparts, &rz_new, &rr);
    if (threadIdx.x == 0) {
        scal[2] = (scal[0] != 0.0) ? rz_new / scal[0] : 0.0;
        scal[0] = rz_new;
        scal[3] = rr;
    }
}

__global__ void __launch_bounds__(kB) k_direction(int64_t n, const double* __restrict__ scal,
                                                  const double* __restrict__ z,
                                                  double* __restrict__ p) {
    const double beta = scal[2];
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
        p[i] = z[i] + beta * p[i];
}

// Setup: dinv = 1/diag (1 on fixed rows), r = b (0 on fixed), z = dinv r, p = z,
// partials of r.z and r.r.
__global__ void __launch_bounds__(kB)
k_setup(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
        const double* __restrict__ vals, const uint8_t* __restrict__ fixed, int64_t n,
        const double* __restrict__ b, double* __restrict__ dinv, double* __restrict__ r,
        double* __restrict__ z, double* __restrict__ p, double* __restrict__ x,
        double* __restrict__ prz, double* __restrict__ prr) {
    __shared__ double red[2][kB / 64];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const bool fi = fixed && fixed[i];
        double d = 1.0;
        if (!fi) {
            d = 0.0;
            for (int k = colptr[i]; k < colptr[i + 1]; ++k)
                if (rowind[k] == i) d += vals[k];
        }
        const double di = (d != 0.0) ? 1.0 / d : 1.0;
        dinv[i] = di;
        const double ri = fi ? 0.0 : b[i];
        x[i] = 0.0;
        r[i] = ri;
        z[i] = ri * di;
        p[i] = ri * di;
        rz += ri * ri * di;
        rr += ri * ri;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        rz += __shfl_xor(rz, off, 64);
        rr += __shfl_xor(rr, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = rz;
        red[1][threadIdx.x >> 6] = rr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int w = 0; w < kB / 64; ++w) {
            a += red[0][w];
            c += red[1][w];
        }
        prz[blockIdx.x] = a;
        prr[blockIdx.x] = c;
    }
}

__global__ void __launch_bounds__(kB) k_setup_scal(const double* __restrict__ prz,
                                                   const double* __restrict__ prr, int nparts,
                                                   double* __restrict__ scal) {
    const double rz = block_sum_fixed(prz, nparts);
    const double rr = block_sum_fixed(prr, nparts);
    if (threadIdx.x == 0) {
        scal[0] = rz;
        scal[1] = 0.0;
        scal[2] = 0.0;
        scal[3] = rr;
    }
}

// y = A_bc x (the Dirichlet-applied operator), no dot product.
__global__ void __launch_bounds__(kB)
k_spmv(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
       const double* __restrict__ vals, const uint8_t* __restrict__ fixed, int mask_mode,
       int64_t n, const double* __restrict__ x, double* __restrict__ y) {
    const int lane = threadIdx.x % kRowLanes;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / kRowLanes);
    for (int64_t i = (int64_t)blockIdx.x * (kB / kRowLanes) + threadIdx.x / kRowLanes; i < n;
         i += rows_per_grid) {
        double s = 0.0;
        for (int k = colptr[i] + lane; k < colptr[i + 1]; k += kRowLanes) s += vals[k] * x[rowind[k]];
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kRowLanes);
        if (lane == 0) y[i] = s;
    }
}

// out = sum_k coef[k] * vals_k (same pattern), up to 4 inputs.
__global__ void __launch_bounds__(kB)
k_axpby4(int64_t nnz, double c0, const double* __restrict__ v0, double c1,
         const double* __restrict__ v1, double c2, const double* __restrict__ v2, double c3,
         const double* __restrict__ v3, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kB) {
        double s = c0 * v0[i];
        if (v1) s += c1 * v1[i];
        if (v2) s += c2 * v2[i];
        if (v3) s += c3 * v3[i];
        out[i] = s;
    }
}

// ---- auxiliary-space (gradient) correction ---------------------------------
// M^-1 r = D^-1 r + G p(Ln) G^T r, with G the edge-node incidence restricted to
// the free nodes and Ln = G^T A G (for a curl-curl operator plus a mass term,
// G^T A G = G^T (mass part) G: the Jacobian's curl part vanishes on
// gradients).  p(Ln) = `sweeps` damped Jacobi sweeps from zero, a fixed SPD
// polynomial, so the preconditioner stays symmetric and CG applies.

__global__ void __launch_bounds__(kB)
k_update_r(int64_t n, const double* __restrict__ scal, const double* __restrict__ p,
           const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
           double* __restrict__ partial_rr) {
    __shared__ double red[kB / 64];
    const double alpha = scal[1];
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        rr += ri * ri;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rr += __shfl_xor(rr, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rr;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < kB / 64; ++w) a += red[w];
        partial_rr[blockIdx.x] = a;
    }
}

// g = G^T r: node v sums +-r over its edges (ref = dof << 1 | v is the higher
// end), kRowLanes lanes per node (fixed butterfly order).  With phi0 given,
// also the first damped Jacobi sweep from zero, phi0 = omega D^-1 g (one
// launch fewer than a kernel of its own).
__global__ void __launch_bounds__(kB)
k_gt(int64_t nv, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nref,
     const double* __restrict__ r, double* __restrict__ g,
     const double* __restrict__ dninv = nullptr, double omega = 0.0,
     double* __restrict__ phi0 = nullptr) {
    const int lane = threadIdx.x % kRowLanes;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / kRowLanes);
    for (int64_t v = (int64_t)blockIdx.x * (kB / kRowLanes) + threadIdx.x / kRowLanes; v < nv;
         v += rows_per_grid) {
        double s = 0.0;
        for (int k = nptr[v] + lane; k < nptr[v + 1]; k += kRowLanes) {
            const int ref = nref[k];
            const double rv = r[ref >> 1];
            s += (ref & 1) ? rv : -rv;
        }
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kRowLanes);
        if (lane == 0) {
            g[v] = s;
            if (phi0) phi0[v] = omega * dninv[v] * s;
        }
    }
}

// phi_out = phi_in + omega D^-1 (g - Ln phi_in), kRowLanes lanes per row.
__global__ void __launch_bounds__(kB)
k_jacobi_sweep(const int32_t* __restrict__ lp, const int32_t* __restrict__ li,
               const double* __restrict__ lv, int64_t nv, const double* __restrict__ dninv,
               const double* __restrict__ g, double omega, const double* __restrict__ phi_in,
               double* __restrict__ phi_out) {
    const int lane = threadIdx.x % kRowLanes;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / kRowLanes);
    for (int64_t v = (int64_t)blockIdx.x * (kB / kRowLanes) + threadIdx.x / kRowLanes; v < nv;
         v += rows_per_grid) {
        double s = 0.0;
        for (int k = lp[v] + lane; k < lp[v + 1]; k += kRowLanes) s += lv[k] * phi_in[li[k]];
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kRowLanes);
        if (lane == 0) phi_out[v] = phi_in[v] + omega * dninv[v] * (g[v] - s);
    }
}

// z = D^-1 r + G phi (0 on fixed rows) and the partials of r.z.
__global__ void __launch_bounds__(kB)
k_z(int64_t n, const int2* __restrict__ en, const uint8_t* __restrict__ fixed,
    const double* __restrict__ dinv, const double* __restrict__ r, const double* __restrict__ phi,
    double* __restrict__ z, double* __restrict__ partial_rz) {
    __shared__ double red[kB / 64];
    double rz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        double zi = 0.0;
        if (!(fixed && fixed[i])) {
            const int2 ab = en[i];
            zi = dinv[i] * r[i];
            if (ab.y >= 0) zi += phi[ab.y];
            if (ab.x >= 0) zi -= phi[ab.x];
        }
        z[i] = zi;
        rz += r[i] * zi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rz += __shfl_xor(rz, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rz;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < kB / 64; ++w) a += red[w];
        partial_rz[blockIdx.x] = a;
    }
}

__global__ void __launch_bounds__(kB)
k_ln_dinv(const int32_t* __restrict__ lp, const int32_t* __restrict__ li,
          const double* __restrict__ lv, int64_t nv, double* __restrict__ dninv) {
    for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < nv; v += (int64_t)gridDim.x * kB) {
        double d = 0.0;
        for (int k = lp[v]; k < lp[v + 1]; ++k)
            if (li[k] == v) d += lv[k];
        dninv[v] = (d > 0.0) ? 1.0 / d : 0.0;
    }
}

__global__ void __launch_bounds__(kB) k_copy(int64_t n, const double* __restrict__ a,
                                             double* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
        b[i] = a[i];
}

// ---- complex symmetric systems (COCG) -----------------------------------------
// A = R + iI with R, I real symmetric on one structural CSC pattern; vectors are
// complex128 (interleaved re, im).  COCG (van der Vorst and Melissen, 1990) is
// CG with the unconjugated bilinear form x^T y, valid for complex symmetric A
// and a complex symmetric preconditioner.  The preconditioner is the real
// auxiliary-space operator above, built from P = R + I (SPD when R is SPD and
// I is positive semi-definite), applied to the real and imaginary parts alike.
// For R, I >= 0 the eigenvalues of P^-1 A lie on {(1 + i t)/(1 + t), t >= 0},
// modulus in [1/sqrt(2), 1], so the iteration count follows that of P.

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    if (d == 0.0) return make_double2(0.0, 0.0);
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Writes a block's (re, im) partial sums: wave butterflies, then wave order.
__device__ inline void block_partial2(double a, double b, double* __restrict__ pa,
                                      double* __restrict__ pb) {
    __shared__ double red[2][kB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, t = 0.0;
        for (int w = 0; w < kB / 64; ++w) {
            s += red[0][w];
            t += red[1][w];
        }
        pa[blockIdx.x] = s;
        if (pb) pb[blockIdx.x] = t;
    }
}

// q = A_bc p (fixed rows and columns removed, unit diagonal) and partials of p^T q.
__global__ void __launch_bounds__(kB)
k_c_spmv_dot(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
             const double* __restrict__ vr, const double* __restrict__ vi,
             const uint8_t* __restrict__ fixed, int64_t n, const double2* __restrict__ x,
             double2* __restrict__ y, double* __restrict__ pre, double* __restrict__ pim) {
    const int lane = threadIdx.x % kRowLanes;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / kRowLanes);
    double dre = 0.0, dim = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * (kB / kRowLanes) + threadIdx.x / kRowLanes; i < n;
         i += rows_per_grid) {
        double sr = 0.0, si = 0.0;
        const bool fi = fixed && fixed[i];
        if (!fi) {
            for (int k = colptr[i] + lane; k < colptr[i + 1]; k += kRowLanes) {
                const int j = rowind[k];
                if (!(fixed && fixed[j])) {
                    const double2 xj = x[j];
                    const double a = vr[k], b = vi[k];
                    sr += a * xj.x - b * xj.y;
                    si += a * xj.y + b * xj.x;
                }
            }
        }
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, kRowLanes);
            si += __shfl_xor(si, off, kRowLanes);
        }
        if (lane == 0) {
            const double2 xi = x[i];
            if (fi) {
                sr = xi.x;
                si = xi.y;
            }
            y[i] = make_double2(sr, si);
            dre += xi.x * sr - xi.y * si;
            dim += xi.x * si + xi.y * sr;
        }
    }
    block_partial2(dre, dim, pre, pim);
}

// dinv = 1/diag(R + I) (1 on fixed rows), r = b (0 on fixed rows), x = 0, partials of |r|^2.
__global__ void __launch_bounds__(kB)
k_c_setup(const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
          const double* __restrict__ vr, const double* __restrict__ vi,
          const uint8_t* __restrict__ fixed, int64_t n, const double2* __restrict__ b,
          double* __restrict__ dinv, double2* __restrict__ r, double2* __restrict__ x,
          double* __restrict__ prr) {
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const bool fi = fixed && fixed[i];
        double d = 1.0;
        if (!fi) {
            d = 0.0;
            for (int k = colptr[i]; k < colptr[i + 1]; ++k)
                if (rowind[k] == i) d += vr[k] + vi[k];
        }
        dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
        const double2 ri = fi ? make_double2(0.0, 0.0) : b[i];
        r[i] = ri;
        x[i] = make_double2(0.0, 0.0);
        rr += ri.x * ri.x + ri.y * ri.y;
    }
    block_partial2(rr, 0.0, prr, nullptr);
}

__global__ void __launch_bounds__(kB)
k_c_gt(int64_t nv, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nref,
       const double2* __restrict__ r, double2* __restrict__ g) {
    for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < nv; v += (int64_t)gridDim.x * kB) {
        double sr = 0.0, si = 0.0;
        for (int k = nptr[v]; k < nptr[v + 1]; ++k) {
            const int ref = nref[k];
            const double2 rv = r[ref >> 1];
            if (ref & 1) {
                sr += rv.x;
                si += rv.y;
            } else {
                sr -= rv.x;
                si -= rv.y;
            }
        }
        g[v] = make_double2(sr, si);
    }
}

__global__ void __launch_bounds__(kB)
k_c_jacobi_first(int64_t nv, const double* __restrict__ dninv, const double2* __restrict__ g,
                 double omega, double2* __restrict__ phi) {
    for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < nv; v += (int64_t)gridDim.x * kB) {
        const double s = omega * dninv[v];
        const double2 gv = g[v];
        phi[v] = make_double2(s * gv.x, s * gv.y);
    }
}

__global__ void __launch_bounds__(kB)
k_c_jacobi_sweep(const int32_t* __restrict__ lp, const int32_t* __restrict__ li,
                 const double* __restrict__ lv, int64_t nv, const double* __restrict__ dninv,
                 const double2* __restrict__ g, double omega, const double2* __restrict__ phi_in,
                 double2* __restrict__ phi_out) {
    const int lane = threadIdx.x % kRowLanes;
    const int64_t rows_per_grid = (int64_t)gridDim.x * (kB / kRowLanes);
    for (int64_t v = (int64_t)blockIdx.x * (kB / kRowLanes) + threadIdx.x / kRowLanes; v < nv;
         v += rows_per_grid) {
        double sr = 0.0, si = 0.0;
        for (int k = lp[v] + lane; k < lp[v + 1]; k += kRowLanes) {
            const double2 pj = phi_in[li[k]];
            sr += lv[k] * pj.x;
            si += lv[k] * pj.y;
        }
#pragma unroll
        for (int off = kRowLanes / 2; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, kRowLanes);
            si += __shfl_xor(si, off, kRowLanes);
        }
        if (lane == 0) {
            const double s = omega * dninv[v];
            const double2 pv = phi_in[v], gv = g[v];
            phi_out[v] = make_double2(pv.x + s * (gv.x - sr), pv.y + s * (gv.y - si));
        }
    }
}

// z = D^-1 r + G phi (0 on fixed rows) and partials of r^T z.
__global__ void __launch_bounds__(kB)
k_c_z(int64_t n, const int2* __restrict__ en, const uint8_t* __restrict__ fixed,
      const double* __restrict__ dinv, const double2* __restrict__ r,
      const double2* __restrict__ phi, double2* __restrict__ z, double* __restrict__ pre,
      double* __restrict__ pim) {
    double dre = 0.0, dim = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        double2 zi = make_double2(0.0, 0.0);
        const double2 ri = r[i];
        if (!(fixed && fixed[i])) {
            zi = make_double2(dinv[i] * ri.x, dinv[i] * ri.y);
            if (phi) {
                const int2 ab = en[i];
                if (ab.y >= 0) {
                    zi.x += phi[ab.y].x;
                    zi.y += phi[ab.y].y;
                }
                if (ab.x >= 0) {
                    zi.x -= phi[ab.x].x;
                    zi.y -= phi[ab.x].y;
                }
            }
        }
        z[i] = zi;
        dre += ri.x * zi.x - ri.y * zi.y;
        dim += ri.x * zi.y + ri.y * zi.x;
    }
    block_partial2(dre, dim, pre, pim);
}

// scal: [0,1] rho, [2,3] alpha, [4,5] beta, [6] |r|^2
__global__ void __launch_bounds__(kB) k_c_alpha(const double* __restrict__ pre,
                                                const double* __restrict__ pim, int nparts,
                                                double* __restrict__ scal) {
    const double mr = block_sum_fixed(pre, nparts);
    const double mi = block_sum_fixed(pim, nparts);
    if (threadIdx.x == 0) {
        const double2 a = cdiv(make_double2(scal[0], scal[1]), make_double2(mr, mi));
        scal[2] = a.x;
        scal[3] = a.y;
    }
}

__global__ void __launch_bounds__(kB)
k_c_update(int64_t n, const double* __restrict__ scal, const double2* __restrict__ p,
           const double2* __restrict__ q, double2* __restrict__ x, double2* __restrict__ r,
           double* __restrict__ prr) {
    const double2 alpha = make_double2(scal[2], scal[3]);
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const double2 ap = cmul(alpha, p[i]);
        const double2 aq = cmul(alpha, q[i]);
        double2 xi = x[i];
        xi.x += ap.x;
        xi.y += ap.y;
        x[i] = xi;
        double2 ri = r[i];
        ri.x -= aq.x;
        ri.y -= aq.y;
        r[i] = ri;
        rr += ri.x * ri.x + ri.y * ri.y;
    }
    block_partial2(rr, 0.0, prr, nullptr);
}

// first == 1: rho = r^T z, p = z.  Otherwise beta = rho_new / rho.  Also |r|^2.
__global__ void __launch_bounds__(kB)
k_c_beta(const double* __restrict__ pre, const double* __restrict__ pim,
         const double* __restrict__ prr, int nparts, int first, double* __restrict__ scal) {
    const double zr = block_sum_fixed(pre, nparts);
    const double zi = block_sum_fixed(pim, nparts);
    const double rr = block_sum_fixed(prr, nparts);
    if (threadIdx.x == 0) {
        const double2 b = first ? make_double2(0.0, 0.0)
                                : cdiv(make_double2(zr, zi), make_double2(scal[0], scal[1]));
        scal[4] = b.x;
        scal[5] = b.y;
        scal[0] = zr;
        scal[1] = zi;
        scal[6] = rr;
    }
}

__global__ void __launch_bounds__(kB) k_c_direction(int64_t n, const double* __restrict__ scal,
                                                    const double2* __restrict__ z,
                                                    double2* __restrict__ p) {
    const double2 beta = make_double2(scal[4], scal[5]);
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const double2 bp = cmul(beta, p[i]);
        const double2 zi = z[i];
        p[i] = make_double2(zi.x + bp.x, zi.y + bp.y);
    }
}

int grid_parts(int64_t n, int per_block) {
    int64_t g = (n + per_block - 1) / per_block;
    if (g > kMaxPartials) g = kMaxPartials;
    if (g < 1) g = 1;
    return (int)g;
}

// Shape of the segmented SpMV: rows per wave, loads in flight and LDS
// products per row.  Default 32 x 4 x 16: a 16 KB block allows 8 waves per
// SIMD, and a group longer than 512 entries takes a second window (983k-tet
// E-core Newton systems: 60 us per product, against 64 us with 20 products
// per row and 68 us with 24).  MPK_SPMV_SHAPE picks 16x4, 32x2, 32x8, 64x4
// (24 products per row), 32x4w20 or 32x4w24 instead (a tuning knob: the
// per-row summation order is the same for every shape).
int spmv_shape() {
    const char* e = std::getenv("MPK_SPMV_SHAPE");
    if (!e) return 0;
    if (!std::strcmp(e, "16x4")) return 1;
    if (!std::strcmp(e, "32x2")) return 2;
    if (!std::strcmp(e, "32x8")) return 3;
    if (!std::strcmp(e, "64x4")) return 4;
    if (!std::strcmp(e, "32x4w20")) return 5;
    if (!std::strcmp(e, "32x4w24")) return 6;
    return 0;
}
int spmv_rows(int shape) { return shape == 1 ? 16 : shape == 4 ? 64 : 32; }
int spmv_grid(int64_t n, int rows) { return grid_parts(n, (kB / 64) * rows); }
void launch_spmv_dot(int shape, int grid, hipStream_t st, const int32_t* colptr,
                     const int32_t* rowind, const double* vals, const uint8_t* fixed, int64_t n,
                     const double* x, double* y, double* partial) {
#define MPK_SPMV(R, U, W)                                                                    \
    hipLaunchKernelGGL((k_spmv_dot<R, U, W>), dim3(grid), dim3(kB), 0, st, colptr, rowind,       \
                       vals, fixed, n, x, y, partial)
    switch (shape) {
        case 1: MPK_SPMV(16, 4, 24); break;
        case 2: MPK_SPMV(32, 2, 24); break;
        case 3: MPK_SPMV(32, 8, 24); break;
        case 4: MPK_SPMV(64, 4, 24); break;
        case 5: MPK_SPMV(32, 4, 20); break;
        case 6: MPK_SPMV(32, 4, 24); break;
        default: MPK_SPMV(32, 4, 16);
    }
#undef MPK_SPMV
}

// Solver stream: a solve runs on a private stream of this thread and device,
// ordered after the caller's stream by an event at the start and before it by
// one at the end.  Its iteration batches are captured once into a HIP graph
// and replayed: on small systems an iteration is a few microsecond-long
// kernels, and one graph launch per batch replaces their individual launches.
// The captured kernels take fixed device pointers (the solver's workspace), so
// every replay runs the same iterations.  MPK_SOLVER_GRAPHS=0 launches the
// kernels one by one (same results).
SideStream* thread_solver_stream(int dev) {
    // created once per thread and device and kept to the end of the process
    static thread_local SideStream* streams[64] = {};
    if (dev < 0 || dev >= 64) return nullptr;
    if (!streams[dev]) streams[dev] = new SideStream();
    return streams[dev];
}

struct SolveStream {
    hipStream_t caller = nullptr, s = nullptr;
    SideStream* side = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graphs = true;
    hipError_t err = hipSuccess;
    explicit SolveStream(hipStream_t c) : caller(c) {
        int dev = 0;
        err = hipGetDevice(&dev);
        if (err != hipSuccess) return;
        side = thread_solver_stream(dev);
        if (!side) {
            err = hipErrorInvalidDevice;
            return;
        }
        err = side->ensure();
        if (err != hipSuccess) return;
        s = side->s;
        err = hipEventRecord(side->fork, caller);
        if (err == hipSuccess) err = hipStreamWaitEvent(s, side->fork, 0);
        const char* e = std::getenv("MPK_SOLVER_GRAPHS");
        graphs = !(e && e[0] == '0');
    }
    ~SolveStream() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (side && s) {
            (void)hipEventRecord(side->join, s);
            (void)hipStreamWaitEvent(caller, side->join, 0);
        }
    }
};

// Launches `batch` iterations (body() launches one) on the solve stream: a
// full batch (`full` iterations) as a replay of the graph captured from the
// first full batch, a shorter one directly.  A capture that fails falls back
// to direct launches for the rest of the solve.
template <class F>
hipError_t run_batch(SolveStream& ss, int batch, int full, F&& body) {
    if (batch == full && ss.graphs) {
        if (!ss.exec) {
            hipError_t e = hipStreamBeginCapture(ss.s, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                for (int k = 0; k < batch; ++k) body();
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(ss.s, &g);
                if (e == hipSuccess) e = hipGraphInstantiate(&ss.exec, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                ss.graphs = false;
                ss.exec = nullptr;
                for (int k = 0; k < batch; ++k) body();
                return hipGetLastError();
            }
        }
        return hipGraphLaunch(ss.exec, ss.s);
    }
    for (int k = 0; k < batch; ++k) body();
    return hipGetLastError();
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_spmv_sym(const int32_t* colptr, const int32_t* rowind, const double* vals, int64_t n,
                 const double* x, double* y, void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 0, MPK_ERR_ARGS, "negative size");
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(colptr && rowind && vals && x && y, MPK_ERR_ARGS, "NULL argument");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_spmv, dim3(grid_parts(n, kB / kRowLanes)), dim3(kB), 0, st, colptr,
                       rowind, vals, nullptr, 0, n, x, y);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_csc_combine(int64_t nnz, double c0, const double* v0, double c1, const double* v1,
                    double c2, const double* v2, double c3, const double* v3, double* out,
                    void* stream) {
    clear_error();
    MPK_REQUIRE(nnz >= 0, MPK_ERR_ARGS, "negative size");
    if (nnz == 0) return MPK_OK;
    MPK_REQUIRE(v0 && out, MPK_ERR_ARGS, "NULL argument");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_axpby4, dim3(grid_parts(nnz, kB)), dim3(kB), 0, st, nnz, c0, v0, c1, v1,
                       c2, v2, c3, v3, out);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_pcg_solve(const int32_t* colptr, const int32_t* rowind, const double* vals, int64_t n,
                  const uint8_t* fixed, const double* b, double* x, double* work,
                  double rel_tol, int max_iter, int check_every, int* iters_out,
                  double* rel_res_out, void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 0 && max_iter >= 0 && iters_out && rel_res_out, MPK_ERR_ARGS,
                "bad argument");
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(colptr && rowind && vals && b && x && work, MPK_ERR_ARGS, "NULL argument");
    if (check_every < 1) check_every = 1;
    SolveStream ss(as_stream(stream));
    MPK_HIP_TRY(ss.err);
    hipStream_t st = ss.s;
    // work layout: dinv, r, z, p, q (n each), partials (3 * kMaxPartials), scal (8)
    double* dinv = work;
    double* r = dinv + n;
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    double* pa = q + n;
    double* pb = pa + kMaxPartials;
    double* pc = pb + kMaxPartials;
    double* scal = pc + kMaxPartials;
    const int g_vec = grid_parts(n, kB);
    const int shape = spmv_shape();
    const int g_row = spmv_grid(n, spmv_rows(shape));
    hipLaunchKernelGGL(k_setup, dim3(g_vec), dim3(kB), 0, st, colptr, rowind, vals, fixed, n, b,
                       dinv, r, z, p, x, pb, pc);
    hipLaunchKernelGGL(k_setup_scal, dim3(1), dim3(kB), 0, st, pb, pc, g_vec, scal);
    MPK_HIP_TRY(hipGetLastError());
    double h[4];
    MPK_HIP_TRY(hipMemcpyAsync(h, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    const double r0 = std::sqrt(h[3]);
    if (r0 == 0.0) return MPK_OK;
    int it = 0;
    double rel = 1.0;
    while (it < max_iter) {
        const int batch = (max_iter - it) < check_every ? (max_iter - it) : check_every;
        MPK_HIP_TRY(run_batch(ss, batch, check_every, [&]() {
            launch_spmv_dot(shape, g_row, st, colptr, rowind, vals, fixed, n, p, q, pa);
            hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kB), 0, st, pa, g_row, scal);
            hipLaunchKernelGGL(k_update, dim3(g_vec), dim3(kB), 0, st, n, scal, p, q, dinv, x, r,
                               z, pb, pc);
            hipLaunchKernelGGL(k_beta, dim3(1), dim3(kB), 0, st, pb, pc, g_vec, scal);
            hipLaunchKernelGGL(k_direction, dim3(g_vec), dim3(kB), 0, st, n, scal, z, p);
        }));
        it += batch;
        MPK_HIP_TRY(hipMemcpyAsync(h, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        MPK_HIP_TRY(hipStreamSynchronize(st));
        rel = std::sqrt(h[3]) / r0;
        if (!(rel > rel_tol)) break;  // also stops on NaN
    }
    *iters_out = it;
    *rel_res_out = rel;
    return MPK_OK;
}

int mpk_pcg_aux_solve(const int32_t* colptr, const int32_t* rowind, const double* vals,
                      int64_t n, const uint8_t* fixed, const mpk_aux_space* aux, const double* b,
                      double* x, double* work, double rel_tol, int max_iter, int check_every,
                      int* iters_out, double* rel_res_out, void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 0 && max_iter >= 0 && iters_out && rel_res_out && aux, MPK_ERR_ARGS,
                "bad argument");
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(colptr && rowind && vals && b && x && work, MPK_ERR_ARGS, "NULL argument");
    const int64_t nv = aux->num_nodes;
    MPK_REQUIRE(nv >= 0 && aux->sweeps >= 1 && aux->edge_nodes &&
                    (nv == 0 || (aux->node_ptr && aux->node_refs && aux->ln_ptr &&
                                 aux->ln_idx && aux->ln_vals)),
                MPK_ERR_ARGS, "incomplete auxiliary space");
    if (check_every < 1) check_every = 1;
    SolveStream ss(as_stream(stream));
    MPK_HIP_TRY(ss.err);
    hipStream_t st = ss.s;
    // work: dinv, r, z, p, q (n each), dninv, g, phi0, phi1 (nv each), partials, scal
    double* dinv = work;
    double* r = dinv + n;
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    double* dninv = q + n;
    double* g = dninv + nv;
    double* phi0 = g + nv;
    double* phi1 = phi0 + nv;
    double* pa = phi1 + nv;
    double* pb = pa + kMaxPartials;
    double* pc = pb + kMaxPartials;
    double* scal = pc + kMaxPartials;
    const int g_vec = grid_parts(n, kB);
    const int shape = spmv_shape();
    const int g_row = spmv_grid(n, spmv_rows(shape));
    const int g_nv = grid_parts(nv, kB);
    const int g_nrow = grid_parts(nv, kB / kRowLanes);
    const int2* en = reinterpret_cast<const int2*>(aux->edge_nodes);
    // z = M^-1 r; partial r.z into pb
    auto precondition = [&]() {
        if (nv > 0) {
            hipLaunchKernelGGL(k_gt, dim3(g_nrow), dim3(kB), 0, st, nv, aux->node_ptr,
                               aux->node_refs, r, g, dninv, aux->omega, phi0);
            double* a = phi0;
            double* c = phi1;
            for (int k = 1; k < aux->sweeps; ++k) {
                hipLaunchKernelGGL(k_jacobi_sweep, dim3(g_nrow), dim3(kB), 0, st, aux->ln_ptr,
                                   aux->ln_idx, aux->ln_vals, nv, dninv, g, aux->omega, a, c);
                double* t = a;
                a = c;
                c = t;
            }
            hipLaunchKernelGGL(k_z, dim3(g_vec), dim3(kB), 0, st, n, en, fixed, dinv, r, a, z,
                               pb);
        } else {
            hipLaunchKernelGGL(k_z, dim3(g_vec), dim3(kB), 0, st, n, en, fixed, dinv, r, phi0, z,
                               pb);
        }
    };
    if (nv > 0)
        hipLaunchKernelGGL(k_ln_dinv, dim3(g_nv), dim3(kB), 0, st, aux->ln_ptr, aux->ln_idx,
                           aux->ln_vals, nv, dninv);
    // setup: dinv, r = b, x = 0 (k_setup also forms a Jacobi z that is replaced)
    hipLaunchKernelGGL(k_setup, dim3(g_vec), dim3(kB), 0, st, colptr, rowind, vals, fixed, n, b,
                       dinv, r, z, p, x, pb, pc);
    precondition();
    hipLaunchKernelGGL(k_setup_scal, dim3(1), dim3(kB), 0, st, pb, pc, g_vec, scal);
    hipLaunchKernelGGL(k_copy, dim3(g_vec), dim3(kB), 0, st, n, z, p);
    MPK_HIP_TRY(hipGetLastError());
    double h[4];
    MPK_HIP_TRY(hipMemcpyAsync(h, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    const double r0 = std::sqrt(h[3]);
    if (r0 == 0.0) return MPK_OK;
    int it = 0;
    double rel = 1.0;
    while (it < max_iter) {
        const int batch = (max_iter - it) < check_every ? (max_iter - it) : check_every;
        MPK_HIP_TRY(run_batch(ss, batch, check_every, [&]() {
            launch_spmv_dot(shape, g_row, st, colptr, rowind, vals, fixed, n, p, q, pa);
            hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kB), 0, st, pa, g_row, scal);
            hipLaunchKernelGGL(k_update_r, dim3(g_vec), dim3(kB), 0, st, n, scal, p, q, x, r, pc);
            precondition();
            hipLaunchKernelGGL(k_beta, dim3(1), dim3(kB), 0, st, pb, pc, g_vec, scal);
            hipLaunchKernelGGL(k_direction, dim3(g_vec), dim3(kB), 0, st, n, scal, z, p);
        }));
        it += batch;
        MPK_HIP_TRY(hipMemcpyAsync(h, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        MPK_HIP_TRY(hipStreamSynchronize(st));
        rel = std::sqrt(h[3]) / r0;
        if (!(rel > rel_tol)) break;  // also stops on NaN
    }
    *iters_out = it;
    *rel_res_out = rel;
    return MPK_OK;
}

int mpk_cocg_aux_work_size(int64_t n, int64_t num_nodes, int64_t* size_out) {
    clear_error();
    MPK_REQUIRE(n >= 0 && num_nodes >= 0 && size_out, MPK_ERR_ARGS, "bad argument");
    *size_out = 9 * (n + 1) + 7 * (num_nodes + 1) + 3 * (int64_t)kMaxPartials + 16;
    return MPK_OK;
}

int mpk_cocg_aux_solve(const int32_t* colptr, const int32_t* rowind, const double* vals_re,
                       const double* vals_im, int64_t n, const uint8_t* fixed,
                       const mpk_aux_space* aux, const double* b, double* x, double* work,
                       double rel_tol, int max_iter, int check_every, int* iters_out,
                       double* rel_res_out, void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 0 && max_iter >= 0 && iters_out && rel_res_out, MPK_ERR_ARGS,
                "bad argument");
    *iters_out = 0;
    *rel_res_out = 0.0;
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(colptr && rowind && vals_re && vals_im && b && x && work, MPK_ERR_ARGS,
                "NULL argument");
    const int64_t nv = aux ? aux->num_nodes : 0;
    MPK_REQUIRE(!aux || (nv >= 0 && aux->sweeps >= 1 && aux->edge_nodes &&
                         (nv == 0 || (aux->node_ptr && aux->node_refs && aux->ln_ptr &&
                                      aux->ln_idx && aux->ln_vals))),
                MPK_ERR_ARGS, "incomplete auxiliary space");
    if (check_every < 1) check_every = 1;
    SolveStream ss(as_stream(stream));
    MPK_HIP_TRY(ss.err);
    hipStream_t st = ss.s;
    // work: dinv (n, padded to even), r, z, p, q (2n each), dninv (nv, padded to even),
    // g, phi0, phi1 (2nv each), partials (3 x kMaxPartials), scal (16)
    double* dinv = work;
    double2* r = reinterpret_cast<double2*>(dinv + ((n + 1) & ~(int64_t)1));
    double2* z = r + n;
    double2* p = z + n;
    double2* q = p + n;
    double* dninv = reinterpret_cast<double*>(q + n);
    double2* g = reinterpret_cast<double2*>(dninv + ((nv + 1) & ~(int64_t)1));
    double2* phi0 = g + nv;
    double2* phi1 = phi0 + nv;
    double* pa = reinterpret_cast<double*>(phi1 + nv);
    double* pb = pa + kMaxPartials;
    double* pc = pb + kMaxPartials;
    double* scal = pc + kMaxPartials;
    MPK_REQUIRE((reinterpret_cast<uintptr_t>(work) & 15) == 0, MPK_ERR_ARGS,
                "work must be 16-byte aligned");
    MPK_REQUIRE((reinterpret_cast<uintptr_t>(b) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(x) & 15) == 0,
                MPK_ERR_ARGS, "b and x must be 16-byte aligned complex128 arrays");
    const double2* b2 = reinterpret_cast<const double2*>(b);
    double2* x2 = reinterpret_cast<double2*>(x);
    const int g_vec = grid_parts(n, kB);
    const int g_row = grid_parts(n, kB / kRowLanes);
    const int g_nv = grid_parts(nv, kB);
    const int g_nrow = grid_parts(nv, kB / kRowLanes);
    const int2* en = aux ? reinterpret_cast<const int2*>(aux->edge_nodes) : nullptr;
    // z = M^-1 r; partials of r^T z into (pa, pb)
    auto precondition = [&]() {
        const double2* phi = nullptr;
        if (nv > 0) {
            hipLaunchKernelGGL(k_c_gt, dim3(g_nv), dim3(kB), 0, st, nv, aux->node_ptr,
                               aux->node_refs, r, g);
            hipLaunchKernelGGL(k_c_jacobi_first, dim3(g_nv), dim3(kB), 0, st, nv, dninv, g,
                               aux->omega, phi0);
            double2* a = phi0;
            double2* c = phi1;
            for (int k = 1; k < aux->sweeps; ++k) {
                hipLaunchKernelGGL(k_c_jacobi_sweep, dim3(g_nrow), dim3(kB), 0, st, aux->ln_ptr,
                                   aux->ln_idx, aux->ln_vals, nv, dninv, g, aux->omega, a, c);
                double2* t = a;
                a = c;
                c = t;
            }
            phi = a;
        }
        hipLaunchKernelGGL(k_c_z, dim3(g_vec), dim3(kB), 0, st, n, en, fixed, dinv, r, phi, z,
                           pa, pb);
    };
    if (nv > 0)
        hipLaunchKernelGGL(k_ln_dinv, dim3(g_nv), dim3(kB), 0, st, aux->ln_ptr, aux->ln_idx,
                           aux->ln_vals, nv, dninv);
    hipLaunchKernelGGL(k_c_setup, dim3(g_vec), dim3(kB), 0, st, colptr, rowind, vals_re, vals_im,
                       fixed, n, b2, dinv, r, x2, pc);
    precondition();
    hipLaunchKernelGGL(k_c_beta, dim3(1), dim3(kB), 0, st, pa, pb, pc, g_vec, 1, scal);
    // p = z (not z + 0 * p: the workspace is uninitialised and may hold NaN)
    hipLaunchKernelGGL(k_copy, dim3(grid_parts(2 * n, kB)), dim3(kB), 0, st, 2 * n,
                       reinterpret_cast<const double*>(z), reinterpret_cast<double*>(p));
    MPK_HIP_TRY(hipGetLastError());
    double h[8];
    MPK_HIP_TRY(hipMemcpyAsync(h, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
    MPK_HIP_TRY(hipStreamSynchronize(st));
    const double r0 = std::sqrt(h[6]);
    if (r0 == 0.0) return MPK_OK;
    int it = 0;
    double rel = 1.0;
    while (it < max_iter) {
        const int batch = (max_iter - it) < check_every ? (max_iter - it) : check_every;
        MPK_HIP_TRY(run_batch(ss, batch, check_every, [&]() {
            hipLaunchKernelGGL(k_c_spmv_dot, dim3(g_row), dim3(kB), 0, st, colptr, rowind,
                               vals_re, vals_im, fixed, n, p, q, pa, pb);
            hipLaunchKernelGGL(k_c_alpha, dim3(1), dim3(kB), 0, st, pa, pb, g_row, scal);
            hipLaunchKernelGGL(k_c_update, dim3(g_vec), dim3(kB), 0, st, n, scal, p, q, x2, r,
                               pc);
            precondition();
            hipLaunchKernelGGL(k_c_beta, dim3(1), dim3(kB), 0, st, pa, pb, pc, g_vec, 0, scal);
            hipLaunchKernelGGL(k_c_direction, dim3(g_vec), dim3(kB), 0, st, n, scal, z, p);
        }));
        it += batch;
        MPK_HIP_TRY(hipMemcpyAsync(h, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
        MPK_HIP_TRY(hipStreamSynchronize(st));
        rel = std::sqrt(h[6]) / r0;
        if (!(rel > rel_tol)) break;  // also stops on NaN
    }
    *iters_out = it;
    *rel_res_out = rel;
    return MPK_OK;
}

int mpk_pcg_aux_work_size(int64_t n, int64_t num_nodes, int64_t* size_out) {
    clear_error();
    MPK_REQUIRE(n >= 0 && num_nodes >= 0 && size_out, MPK_ERR_ARGS, "bad argument");
    *size_out = 5 * n + 4 * num_nodes + 3 * (int64_t)kMaxPartials + 8;
    return MPK_OK;
}

int mpk_pcg_work_size(int64_t n, int64_t* size_out) {
    clear_error();
    MPK_REQUIRE(n >= 0 && size_out, MPK_ERR_ARGS, "bad argument");
    *size_out = 5 * n + 3 * (int64_t)kMaxPartials + 8;
    return MPK_OK;
}

}  // extern "C"

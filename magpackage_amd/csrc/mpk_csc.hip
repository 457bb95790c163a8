// Column-owner CSC assembly (the hot path) on gfx950: launchers and C ABI.
//
// Design: one wave64 block owns 64 consecutive CSC columns (edge DOFs), one
// lane per column.  The block's contiguous value range
// [colptr[j0], colptr[j0+64]) is accumulated in LDS; each lane walks the
// incidence list of its column (the elements around the edge), computes the
// element's column of the local block in registers and adds the 6 values at
// their precomputed in-column positions.  A lane only ever touches its own
// column, so there are no atomics and the summation order per CSC entry is
// fixed.  The finished tile leaves the CU as contiguous, coalesced stores.
// HBM traffic per tet is the mesh inputs, the plan's incidence records and
// the CSC values written once; the 36 COO triplets (864 B/tet in the
// reference's MEX output) are never materialised.
//
//   FP64_EXACT  mpk_csc_exact.hpp: reference operation order, element order,
//               bit-identical to MEX kernel + sparse() in triplet order.
//   FP64        mpk_csc_frame.hpp: closed forms with FMA, <= 1e-12 relative.
//   FP32_LOCAL  mpk_csc_frame.hpp: FP32 block entries and element state,
//               FP64 geometry and accumulation, <= 1e-6 relative.
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "mpk_csc_exact.hpp"
#include "mpk_csc_frame.hpp"

namespace mpk {
namespace {

using namespace csc;

__global__ void k_pad_coords(const double* __restrict__ P, int64_t nn, double4* __restrict__ P4) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    P4[i] = make_double4(P[3 * i], P[3 * i + 1], P[3 * i + 2], 0.0);
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
    return TileCtx{plan->colptr,   plan->inc_ptr,   plan->inc_slot,   plan->inc_pos,
                   plan->inc_frame, plan->inc_ring, plan->frame_base, plan->T,
                   plan->P4,        plan->rowind,   plan->n,          plan->inc_ring8,
                   plan->col_head};
}

template <class Op>
int lds_bytes(const mpk_plan* plan, Op& op, size_t acc_size, size_t extra, size_t* lds) {
    const size_t acc_bytes = ((size_t)plan->max_tile_nnz * acc_size + 15) & ~size_t(15);
    *lds = acc_bytes + extra;
    MPK_REQUIRE(*lds <= (size_t)kMaxLdsBytes, MPK_ERR_LIMIT,
                "tile of %lld entries needs %zu B of LDS (limit %d)",
                (long long)plan->max_tile_nnz, *lds, kMaxLdsBytes);
    op.acc_bytes_max = (int)acc_bytes;
    return MPK_OK;
}

int load_coords(const mpk_plan* plan, const double* P, hipStream_t st);

// Element state for the Newton kernels, one lane per element: B = J v / det
// with v = sum_i a_i s_i c_i over the constant reference curls c_i
// (nedelec_tet_p1.m:48; the Piola map of assemble_jacobian_kernel_mex.cpp:
// 94-113), then (nu, dnu) = evaluate_nu_derivative(|B|^2, tag)
// (MexMaterialUtils.hpp:24-87) with the B-H tables staged in LDS.
// The FP64 state (StateView) is 48 B per element; the FP32-local one
// (StateView32) rounds the same values to FP32, 24 B per element.
constexpr int kStateBlock = 256;

// put_state stores the state and returns (dnu, sigma) as the assembly
// kernels will read them back.
__device__ inline double2 put_state(double4* s0, double2* s1, int64_t e, double bx, double by,
                                    double bz, double nu, double dnu, double sg) {
    s0[e] = make_double4(bx, by, bz, nu);
    s1[e] = make_double2(dnu, sg);
    return make_double2(dnu, sg);
}
__device__ inline double2 put_state(float4* s0, float2* s1, int64_t e, double bx, double by,
                                    double bz, double nu, double dnu, double sg) {
    s0[e] = make_float4((float)bx, (float)by, (float)bz, (float)nu);
    const float2 v = make_float2((float)dnu, (float)sg);
    s1[e] = v;
    return make_double2(v.x, v.y);
}

// Clears the tile class flags and the class-1 tile count.
__global__ void k_clear_tiles(int32_t* __restrict__ flag, int64_t ntiles,
                              int32_t* __restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntiles) flag[i] = 0;
    if (i == 0) *count = 0;
}

// Lists the class-1 tiles (flag == epoch): ascending within each block of 256
// tiles, the blocks' ranges appended in arrival order (the tiles are
// independent, so the list order changes no result).
constexpr int kListBlock = 256;
__global__ void __launch_bounds__(kListBlock)
k_tile_list(const int32_t* __restrict__ flag, int64_t ntiles, int32_t* __restrict__ list,
            int32_t* __restrict__ count, int32_t epoch) {
    __shared__ int wsum[kListBlock / 64];
    __shared__ int base;
    const int64_t t = (int64_t)blockIdx.x * kListBlock + threadIdx.x;
    const bool on = t < ntiles && flag[t] == epoch;
    const unsigned long long m = __ballot(on);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < kListBlock / 64; ++i) tot += wsum[i];
        base = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    int off = base;
    for (int i = 0; i < w; ++i) off += wsum[i];
    if (on) list[off + before] = (int32_t)t;
}

// Elements [e_begin, ne) of the mesh.  kStateElems elements per thread:
// every independent load of all of them
// (connectivity, cell DOFs, tags, sigma) goes out first, then the node and
// solution gathers of all of them, then the arithmetic, so each wave keeps
// twice the loads in flight of a one-element thread.
constexpr int kStateElems = 2;

template <class S0, class S1>
__global__ void __launch_bounds__(kStateBlock)
k_elem_state(const int4* __restrict__ T, const int32_t* __restrict__ cell,
             const double4* __restrict__ P4, int64_t e_begin, int64_t ne,
             const double* __restrict__ sol,
             const int32_t* __restrict__ tags, const double* __restrict__ sigma, MaterialView mat,
             int ntags, int nbrk, int ncoe, S0* __restrict__ s0, S1* __restrict__ s1,
             int32_t* __restrict__ tile_flag, int flag_sigma, int flag_dnu, int32_t epoch,
             int32_t* __restrict__ count, double* __restrict__ mz, int64_t mz_n) {
    __shared__ double tab[kMatLds];
    // Zeroes mz[0, mz_n) (M of the Newton set; its tiles with an M term are
    // rewritten by the pass that runs them) with streaming 16-byte stores,
    // grid-stride over the pairs: fire-and-forget stores issued while this
    // gather-bound kernel waits on its loads.
    if (mz) {
        const int64_t a = (int64_t)((reinterpret_cast<uintptr_t>(mz) >> 3) & 1u);
        const int64_t npair = (mz_n - a) >> 1;
        double2* p2 = reinterpret_cast<double2*>(mz + a);
        const int64_t nth = (int64_t)gridDim.x * kStateBlock;
        for (int64_t q = (int64_t)blockIdx.x * kStateBlock + threadIdx.x; q < npair; q += nth)
            st_out(p2 + q, make_double2(0.0, 0.0));
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (a && mz_n > 0) st_out(mz, 0.0);
            if (mz_n > a && ((mz_n - a) & 1)) st_out(mz + mz_n - 1, 0.0);
        }
    }
    // the class-1 tile count starts at 0 (no kernel reads it before this one ends)
    if (count && blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
    const double* brk = mat.brk;
    const double* coe = mat.coe;
    if (nbrk + ncoe <= kMatLds) {
        for (int i = threadIdx.x; i < nbrk; i += kStateBlock) tab[i] = mat.brk[i];
        for (int i = threadIdx.x; i < ncoe; i += kStateBlock) tab[nbrk + i] = mat.coe[i];
        brk = tab;
        coe = tab + nbrk;
    }
    __syncthreads();
    // XCD-contiguous element ranges: blocks are dealt round-robin over the
    // 8 XCDs, and neighbouring elements share nodes and edges, so each XCD
    // takes one contiguous range and its L2 keeps the shared data
    const int64_t e0 =
        e_begin + xcd_tile(blockIdx.x, gridDim.x) * (kStateBlock * kStateElems) + threadIdx.x;
    if (e0 >= ne) return;
    int64_t e[kStateElems];
    bool ok[kStateElems];
    int4 t[kStateElems];
    int2 c[kStateElems][3];
    int tg[kStateElems];
    double sg[kStateElems];
#pragma unroll
    for (int k = 0; k < kStateElems; ++k) {
        const int64_t ek = e0 + (int64_t)k * kStateBlock;
        ok[k] = ek < ne;
        e[k] = ok[k] ? ek : e0;  // a valid element for the loads
        t[k] = T[e[k]];
        const int2* cr = reinterpret_cast<const int2*>(cell + 6 * e[k]);
        c[k][0] = cr[0];
        c[k][1] = cr[1];
        c[k][2] = cr[2];
        tg[k] = tags[e[k]];
        sg[k] = sigma ? sigma[e[k]] : 1.0;
    }
    Vec3 x[kStateElems][4];
    double al[kStateElems][6];
#pragma unroll
    for (int k = 0; k < kStateElems; ++k) {
        x[k][0] = ld_node(P4, t[k].x);
        x[k][1] = ld_node(P4, t[k].y);
        x[k][2] = ld_node(P4, t[k].z);
        x[k][3] = ld_node(P4, t[k].w);
        const int nd[4] = {t[k].x, t[k].y, t[k].z, t[k].w};
        const int dof[6] = {c[k][0].x, c[k][0].y, c[k][1].x, c[k][1].y, c[k][2].x, c[k][2].y};
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double sgn = nd[edge_a(i)] <= nd[edge_b(i)] ? 1.0 : -1.0;
            al[k][i] = dof[i] >= 0 ? sol[dof[i]] * sgn : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < kStateElems; ++k) {
        if (!ok[k]) continue;
        const Vec3 d1 = sub(x[k][1], x[k][0]), d2 = sub(x[k][2], x[k][0]),
                   d3 = sub(x[k][3], x[k][0]);
        const double det = dot(d1, cross(d2, d3));
        // v = sum_i al_i c_i, c = (0,-2,2) (2,0,-2) (-2,2,0) (0,0,2) (0,-2,0) (2,0,0)
        const double* a = al[k];
        const double vx = 2.0 * (a[1] - a[2] + a[5]);
        const double vy = 2.0 * (a[2] - a[0] - a[4]);
        const double vz = 2.0 * (a[0] - a[1] + a[3]);
        const double inv = 1.0 / det;
        const double bx = (d1.x * vx + d2.x * vy + d3.x * vz) * inv;
        const double by = (d1.y * vx + d2.y * vy + d3.y * vz) * inv;
        const double bz = (d1.z * vx + d2.z * vy + d3.z * vz) * inv;
        // tags are validated by the host layer; clamp so a bad tag cannot
        // read outside the material table
        const int tag = tg[k] < 1 ? 1 : (tg[k] > ntags ? ntags : tg[k]);
        double nu, dnu;
        eval_nu(bx * bx + by * by + bz * bz, tag, mat, brk, coe, nu, dnu);
        const double2 ds = put_state(s0, s1, e[k], bx, by, bz, nu, dnu, sg[k]);
        // Tile classes for the split Newton set: an element with a D term
        // (dnu != 0, the test FrameNewton applies) or an M term (|sigma| >
        // 1e-12 on the caller's sigma, the test of the sigma-mass op) marks the
        // column tiles of its 6 DOFs with the call's epoch.  Every writer stores
        // the same value, so plain stores need no atomics.
        if (tile_flag && ((flag_dnu && ds.x != 0.0) || (flag_sigma && fabs(sg[k]) > 1e-12))) {
            const int dof[6] = {c[k][0].x, c[k][0].y, c[k][1].x, c[k][1].y, c[k][2].x, c[k][2].y};
#pragma unroll
            for (int i = 0; i < 6; ++i)
                if (dof[i] >= 0) tile_flag[dof[i] / kTileCols] = epoch;
        }
    }
}

// Launch context of the element-state kernel over element ranges, with the
// tile-class clear before and the class-1 tile list after.
template <class SV>
struct ElemStateRun {
    const mpk_plan* plan = nullptr;
    const double* sol = nullptr;
    const int32_t* tags = nullptr;
    const double* sigma = nullptr;
    MaterialView mv{};
    int ntags = 0, nbrk = 0, ncoe = 0;
    typename SV::S0* s0 = nullptr;
    typename SV::S1* s1 = nullptr;
    int32_t* flag = nullptr;
    int flag_sigma = 0, flag_dnu = 0;
    int32_t epoch = 1;
    double* mz = nullptr;  // zeroed by the state kernel (the Newton set's M)
    int64_t mz_n = 0;
    int64_t ntiles = 0;
    hipError_t clear(hipStream_t st) const {
        if (flag)
            hipLaunchKernelGGL(k_clear_tiles, dim3(grid_for(ntiles, 256)), dim3(256), 0, st,
                               flag, ntiles, flag + 2 * ntiles);
        return hipGetLastError();
    }
    hipError_t run(int64_t b, int64_t e, hipStream_t st) const {
        if (e > b)
            hipLaunchKernelGGL((k_elem_state<typename SV::S0, typename SV::S1>),
                               dim3(grid_for(e - b, kStateBlock * kStateElems)),
                               dim3(kStateBlock), 0, st, plan->T, plan->cell, plan->P4, b, e, sol,
                               tags, sigma, mv, ntags, nbrk, ncoe, s0, s1, flag, flag_sigma,
                               flag_dnu, epoch, flag ? flag + 2 * ntiles : nullptr,
                               b == 0 ? mz : nullptr, mz_n);
        return hipGetLastError();
    }
    hipError_t list(hipStream_t st) const {
        if (flag)
            hipLaunchKernelGGL(k_tile_list, dim3(grid_for(ntiles, kListBlock)), dim3(kListBlock),
                               0, st, flag, ntiles, flag + ntiles, flag + 2 * ntiles, epoch);
        return hipGetLastError();
    }
};

// Loads P (or checks that the plan holds coordinates), makes sure the state
// buffer exists and fills in the launch context.  The buffer is the plan's,
// allocated on the first call, unless state_buf (ne x sizeof(S0) + ne x
// sizeof(S1) bytes) is given.  With tile_flag ([2 ntiles + 1] int32: class
// flags, class-1 tile list, class-1 count), the tile classes and the list of
// class-1 tiles are written too: flag_sigma / flag_dnu say which terms (M / D)
// classify a tile.  run->plan stays NULL for a plan without elements.
template <class SV>
int elem_state_prepare(const mpk_plan* plan, const double* P, const int32_t* tags,
                       const double* sol, const double* sigma, const mpk_material* mat,
                       hipStream_t st, SV* sv, int32_t* tile_flag, bool flag_sigma,
                       bool flag_dnu, void* state_buf, ElemStateRun<SV>* run) {
    using S0 = typename SV::S0;
    using S1 = typename SV::S1;
    if (P) {
        if (int rc = load_coords(plan, P, st)) return rc;
    } else {
        MPK_REQUIRE(plan->coords_loaded, MPK_ERR_ARGS,
                    "P is NULL and the plan holds no coordinates (mpk_plan_set_coords)");
    }
    if (plan->ne == 0) return MPK_OK;
    MPK_REQUIRE(plan->cell, MPK_ERR_ARGS, "plan without cell DOFs");
    S0* s0 = static_cast<S0*>(state_buf);
    if (!s0) {
        // sized for the FP64 state; the FP32 one uses the first half
        if (!plan->buf_state.p) MPK_HIP_TRY(plan->buf_state.alloc((size_t)plan->ne * 48));
        s0 = plan->buf_state.as<S0>();
    }
    S1* s1 = reinterpret_cast<S1*>(s0 + plan->ne);
    run->plan = plan;
    run->sol = sol;
    run->tags = tags;
    run->sigma = sigma;
    run->mv = MaterialView{mat->linear_nu, mat->is_nonlinear, mat->max_bsq,    mat->breaks,
                           mat->coefs,     mat->spline_start, mat->spline_count, mat->coef_start};
    run->ntags = (int)mat->num_tags;
    run->nbrk = (int)mat->num_breaks;
    run->ncoe = (int)mat->num_coefs;
    run->s0 = s0;
    run->s1 = s1;
    run->flag = tile_flag;
    run->flag_sigma = (int)flag_sigma;
    run->flag_dnu = (int)flag_dnu;
    run->ntiles = (plan->n + kTileCols - 1) / kTileCols;
    *sv = SV{s0, s1};
    return MPK_OK;
}

// The element state of the whole mesh on one stream: clear, state, list.
template <class SV>
int elem_state(const mpk_plan* plan, const double* P, const int32_t* tags, const double* sol,
               const double* sigma, const mpk_material* mat, hipStream_t st, SV* sv,
               int32_t* tile_flag = nullptr, bool flag_sigma = false, bool flag_dnu = false,
               void* state_buf = nullptr) {
    ElemStateRun<SV> r;
    if (int rc = elem_state_prepare(plan, P, tags, sol, sigma, mat, st, sv, tile_flag,
                                    flag_sigma, flag_dnu, state_buf, &r))
        return rc;
    if (!r.plan) return MPK_OK;
    MPK_HIP_TRY(r.clear(st));
    MPK_HIP_TRY(r.run(0, plan->ne, st));
    MPK_HIP_TRY(r.list(st));
    return MPK_OK;
}

template <class Op>
int launch_exact(const mpk_plan* plan, const double* P, Op op, size_t acc_size, size_t extra,
                 hipStream_t st) {
    if (plan->n == 0) return MPK_OK;
    MPK_REQUIRE(P, MPK_ERR_ARGS, "FP64_EXACT mode needs the coordinate array P");
    size_t lds = 0;
    if (int rc = lds_bytes(plan, op, acc_size, extra, &lds)) return rc;
    hipLaunchKernelGGL(k_csc_exact<Op>, dim3(grid_for(plan->n, kW)), dim3(kW), lds, st,
                       tile_ctx(plan), op);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int load_coords(const mpk_plan* plan, const double* P, hipStream_t st) {
    if (plan->nn > 0 && plan->P4) {
        hipLaunchKernelGGL(k_pad_coords, dim3(grid_for(plan->nn, 256)), dim3(256), 0, st, P,
                           plan->nn, plan->P4);
        MPK_HIP_TRY(hipGetLastError());
    }
    plan->coords_loaded = true;
    return MPK_OK;
}

// Incidences per gather round of the ring driver: the op's kRing (else
// kRingBatch) unless the environment sets MPK_RING_BATCH to 2, 3, 4, 6 or 8 (a tuning knob: results
// are identical for every value, the per-entry summation order is the same).
template <class Op, class = void>
struct RingDefault {
    static constexpr int value = kRingBatch;
};
template <class Op>
struct RingDefault<Op, std::void_t<decltype(Op::kRing)>> {
    static constexpr int value = Op::kRing;
};
template <class Op>
int ring_batch() {
    const char* s = std::getenv("MPK_RING_BATCH");
    if (!s) return RingDefault<Op>::value;
    const int v = std::atoi(s);
    return (v == 2 || v == 3 || v == 4 || v == 6 || v == 8) ? v : RingDefault<Op>::value;
}

// Closed-form kernels read the plan's 32-byte-padded coordinate copy.  With P
// given it is refreshed from P first (stream-ordered, so a P changed between
// calls is picked up); with P == NULL the copy made by the last
// mpk_plan_set_coords (or the last call with P) is used.  The ring driver runs
// when the plan has ring records, unless the environment sets
// MPK_CSC_DRIVER=element (element-order driver, for tests and timing).
bool ring_driver(const mpk_plan* plan) {
    const char* drv = std::getenv("MPK_CSC_DRIVER");
    return plan->ring && !(drv && std::strcmp(drv, "element") == 0);
}

// Launches of the ring kernels for a gather batch chosen at run time.
template <class Op, bool kC>
void launch_batch_k(int kb, unsigned grid, size_t lds, hipStream_t st, const TileCtx& tc,
                    const int2* ends, const Op& op, const int32_t* sel, int32_t want,
                    int32_t epoch) {
    switch (kb) {
        case 2:
            hipLaunchKernelGGL((k_csc_ring_batch<Op, 2, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, sel, want, epoch);
            break;
        case 4:
            hipLaunchKernelGGL((k_csc_ring_batch<Op, 4, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, sel, want, epoch);
            break;
        case 6:
            hipLaunchKernelGGL((k_csc_ring_batch<Op, 6, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, sel, want, epoch);
            break;
        case 8:
            hipLaunchKernelGGL((k_csc_ring_batch<Op, 8, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, sel, want, epoch);
            break;
        default:
            hipLaunchKernelGGL((k_csc_ring_batch<Op, 3, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, sel, want, epoch);
    }
}
template <class Op, bool kC>
void launch_list_k(int kb, unsigned grid, size_t lds, hipStream_t st, const TileCtx& tc,
                   const int2* ends, const Op& op, const int32_t* list, const int32_t* count) {
    switch (kb) {
        case 2:
            hipLaunchKernelGGL((k_csc_ring_list<Op, 2, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, list, count);
            break;
        case 4:
            hipLaunchKernelGGL((k_csc_ring_list<Op, 4, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, list, count);
            break;
        case 6:
            hipLaunchKernelGGL((k_csc_ring_list<Op, 6, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, list, count);
            break;
        case 8:
            hipLaunchKernelGGL((k_csc_ring_list<Op, 8, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, list, count);
            break;
        default:
            hipLaunchKernelGGL((k_csc_ring_list<Op, 3, kC>), dim3(grid), dim3(kW), lds, st, tc,
                               ends, op, list, count);
    }
}

// Grid of a persistent list launch: as many blocks as fit on the device at
// `lds` bytes each (at most max_per_cu per CU), no more than the plan's tiles.
int list_grid(const mpk_plan* plan, size_t lds, size_t max_per_cu, unsigned* grid) {
    if (plan->num_cus == 0) {
        int cus = 0;
        MPK_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                          plan->device));
        plan->num_cus = cus > 0 ? cus : 256;
    }
    constexpr size_t kLdsPerCu = 160 * 1024;
    size_t per_cu = lds ? kLdsPerCu / lds : 16;
    per_cu = per_cu < 1 ? 1 : (per_cu > max_per_cu ? max_per_cu : per_cu);
    const int64_t ntiles = (plan->n + kTileCols - 1) / kTileCols;
    // the kernels split the list over the 8 XCDs by block index: the grid is
    // a multiple of 8 so that every XCD's range has blocks
    int64_t g = (int64_t)plan->num_cus * (int64_t)per_cu;
    g = g < ntiles ? g : ntiles;
    g = (g + 7) / 8 * 8;
    *grid = (unsigned)(g > 8 ? g : 8);
    return MPK_OK;
}

// Persistent launch of the ring driver over the tiles of a device list
// (list[0, *count)): as many blocks as fit on the device at the op's LDS size.
template <class Op>
int launch_ring_list(const mpk_plan* plan, Op op, size_t acc_size, hipStream_t st,
                     const int32_t* list, const int32_t* count, size_t max_per_cu = 16,
                     int kb = 0) {
    if (kb == 0) kb = ring_batch<Op>();
    size_t lds = 0;
    if (int rc = lds_bytes(plan, op, acc_size, 0, &lds)) return rc;
    unsigned grid = 0;
    if (int rc = list_grid(plan, lds, max_per_cu, &grid)) return rc;
    const TileCtx tc = tile_ctx(plan);
    if (plan->ring8)
        launch_list_k<Op, true>(kb, grid, lds, st, tc, plan->col_ends, op, list, count);
    else
        launch_list_k<Op, false>(kb, grid, lds, st, tc, plan->col_ends, op, list, count);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

// The two-op persistent list launch (k_csc_ring_list2): op1 then op2 on each
// listed tile, each with its own gather batch; per-CU block count from the
// larger LDS reservation.
template <class Op1, class Op2>
int launch_ring_list2(const mpk_plan* plan, Op1 op1, size_t acc1, Op2 op2, size_t acc2,
                      hipStream_t st, const int32_t* list, const int32_t* count,
                      size_t max_per_cu, int kb1) {
    size_t l1 = 0, l2 = 0;
    if (int rc = lds_bytes(plan, op1, acc1, 0, &l1)) return rc;
    if (int rc = lds_bytes(plan, op2, acc2, 0, &l2)) return rc;
    const size_t lds = l1 > l2 ? l1 : l2;
    unsigned grid = 0;
    if (int rc = list_grid(plan, lds, max_per_cu, &grid)) return rc;
    const TileCtx tc = tile_ctx(plan);
    constexpr int kB2 = RingDefault<Op2>::value;
    if (kb1 == 6) {
        if (plan->ring8)
            hipLaunchKernelGGL((k_csc_ring_list2<Op1, Op2, 6, kB2, true>), dim3(grid), dim3(kW),
                               lds, st, tc, plan->col_ends, op1, op2, list, count);
        else
            hipLaunchKernelGGL((k_csc_ring_list2<Op1, Op2, 6, kB2, false>), dim3(grid), dim3(kW),
                               lds, st, tc, plan->col_ends, op1, op2, list, count);
    } else {
        if (plan->ring8)
            hipLaunchKernelGGL((k_csc_ring_list2<Op1, Op2, 4, kB2, true>), dim3(grid), dim3(kW),
                               lds, st, tc, plan->col_ends, op1, op2, list, count);
        else
            hipLaunchKernelGGL((k_csc_ring_list2<Op1, Op2, 4, kB2, false>), dim3(grid), dim3(kW),
                               lds, st, tc, plan->col_ends, op1, op2, list, count);
    }
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

// sel / want: ring driver only, run the tiles whose class flag equals want.
template <class Op>
int launch_frame(const mpk_plan* plan, const double* P, Op op, size_t acc_size, size_t extra,
                 hipStream_t st, const int32_t* sel = nullptr, int32_t want = 0,
                 int32_t epoch = 1) {
    if (plan->n == 0) return MPK_OK;
    size_t lds = 0;
    if (int rc = lds_bytes(plan, op, acc_size, extra, &lds)) return rc;
    if (P) {
        if (int rc = load_coords(plan, P, st)) return rc;
    } else {
        MPK_REQUIRE(plan->coords_loaded, MPK_ERR_ARGS,
                    "P is NULL and the plan holds no coordinates (mpk_plan_set_coords)");
    }
    const unsigned grid = grid_for(plan->n, kW);
    if (ring_driver(plan)) {
        const TileCtx tc = tile_ctx(plan);
        if (plan->ring8)
            launch_batch_k<Op, true>(ring_batch<Op>(), grid, lds, st, tc, plan->col_ends, op, sel,
                                     want, epoch);
        else
            launch_batch_k<Op, false>(ring_batch<Op>(), grid, lds, st, tc, plan->col_ends, op,
                                      sel, want, epoch);
    } else {
        MPK_REQUIRE(!sel, MPK_ERR_ARGS, "tile selection needs the ring driver");
        hipLaunchKernelGGL(k_csc_frame<Op>, dim3(grid), dim3(kW), lds, st, tile_ctx(plan), op);
    }
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

// Persistent blocks per CU of the Newton set's listed pass when it runs beside
// the class-0 pass.
constexpr size_t kSidePerCu = 2;

// The fused Newton set.  Ring driver: the element-state kernel also classes
// the column tiles; tiles that no M or D term reaches run FrameNewtonLin (one
// LDS plane), the others FrameNewton (two or three planes), in two launches.
// Each tile is written by exactly one of them.
// With scratch_state / scratch_flags given (the allocation-free entry point),
// the element state and the tile classes live there instead of in the plan's
// buffers, and the two passes run one after the other on the caller's stream
// (the side-by-side schedule would create the plan's side stream).
template <int kPrec, class SV>
int newton_set(const mpk_plan* plan, const double* P, const int32_t* tags, const double* sol_a,
               const mpk_material* mat, const double* sigma, const CscQuad& qk,
               const CscQuad& qm, double* k_vals, double* m_vals, double* j_vals,
               double* residual, hipStream_t st, void* scratch_state = nullptr,
               int32_t* scratch_flags = nullptr) {
    const bool with_m = m_vals != nullptr;
    const bool split = plan->n > 0 && ring_driver(plan);
    const bool noalloc = scratch_state != nullptr;
    int32_t* flags = nullptr;
    const int64_t ntiles = (plan->n + kTileCols - 1) / kTileCols;
    // Tile classes: the element state writes this call's epoch into the flag
    // of every class-1 tile.  The plan's flags are cleared once, when they are
    // allocated, and each call takes the next epoch (cleared again only when
    // the epoch would wrap); the allocation-free entry point clears its
    // scratch flags on every call and uses epoch 1.
    int32_t epoch = 1;
    bool clear_flags = false;
    if (split) {
        if (noalloc) {
            flags = scratch_flags;
            clear_flags = true;
        } else {
            if (!plan->buf_tile_flag.p) {
                MPK_HIP_TRY(
                    plan->buf_tile_flag.alloc((size_t)(2 * ntiles + 1) * sizeof(int32_t)));
                plan->tile_epoch = 0;
                clear_flags = true;
            }
            if (plan->tile_epoch >= 0x7ffffff0) {
                plan->tile_epoch = 0;
                clear_flags = true;
            }
            epoch = ++plan->tile_epoch;
            flags = plan->buf_tile_flag.as<int32_t>();
        }
    }
    // M is zero on the tiles no M term reaches.  The element-state kernel
    // zeroes the whole M array with streaming stores while it waits on its
    // gathers; the class-0 pass then stores K and J only, and the listed
    // passes rewrite their tiles' M (4M-tet Newton workload, FP32_LOCAL: a
    // fill kernel of its own before the element state, 104 + 111 us, against
    // 208 us for the two together).  MPK_NEWTON_MFILL=0 stores the zeros from
    // the class-0 pass instead; the allocation-free entry point always does.
    bool mfill = split && with_m && !noalloc && plan->nnz > 0;
    if (const char* e = std::getenv("MPK_NEWTON_MFILL")) mfill = mfill && std::atoi(e) != 0;
    // The classes depend on sigma and dnu, not on which outputs are requested
    // (a call for a subset of the outputs runs each tile with the same op as
    // the full call).  Without sigma the state holds sigma = 1: M then reaches
    // every tile.
    SV sv{};
    ElemStateRun<SV> esr;
    if (int rc = elem_state_prepare(plan, P, tags, sol_a, sigma, mat, st, &sv, flags,
                                    with_m || sigma != nullptr, true, scratch_state, &esr))
        return rc;
    esr.epoch = epoch;
    // without elements there is no element-state launch to zero M or reset
    // the count
    if (split && !esr.plan) {
        clear_flags = true;
        mfill = false;
    }
    if (mfill) {
        esr.mz = m_vals;
        esr.mz_n = plan->nnz;
    }
    // Schedule, by measurement on the 4M-tet Newton workload: the listed
    // tiles run beside the class-0 pass, on the plan's side stream with a few
    // persistent blocks per CU.  FP64 state: one three-plane pass, 2 blocks
    // per CU (0.90 -> 0.84 ms at the time; its split into a two-plane pass
    // and a mass pass measured 0.705-0.708 against 0.699 ms).  FP32 state:
    // the two-plane pass and the mass pass, 4 blocks per CU (0.657 ms; the
    // same two passes after the class-0 pass 0.663-0.664 ms, the three-plane
    // pass after it 0.663-0.665 ms).  The side passes get SIMD slots while
    // the one-plane pass runs only where their waves' registers fit beside
    // the one-plane waves: FP32, one side wave beside two one-plane waves
    // (192 + 2 x 160 of 512 VGPRs); FP64, beside one.  A build with 168 for
    // the one-plane pass starved them until its grid drained (0.72 ms);
    // tests/test_kernel_resources.py checks these budgets on the built
    // library.  MPK_NEWTON_OVERLAP=0/1, MPK_NEWTON_SPLIT_M=0/1,
    // MPK_LIST_BLOCKS_PER_CU (1..16) and MPK_LIST_RING_BATCH (incidences per
    // gather round of the side passes: 2, 3, 4, 6, 8; 4 and 3 measured
    // within 0.4% of the default) override.
    bool overlap = true;
    if (const char* e = std::getenv("MPK_NEWTON_OVERLAP")) overlap = std::atoi(e) != 0;
    bool split_m = kPrec == MPK_PREC_FP32_LOCAL;
    if (const char* e = std::getenv("MPK_NEWTON_SPLIT_M")) split_m = std::atoi(e) != 0;
    size_t per_cu = overlap ? (kPrec == MPK_PREC_FP64 ? kSidePerCu : 4) : 16;
    if (const char* e = std::getenv("MPK_LIST_BLOCKS_PER_CU")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 16) per_cu = (size_t)v;
    }
    // FP32 listed tiles: the two-plane op and the mass op one after the other
    // on each tile in one persistent launch (k_csc_ring_list2, the mass op's
    // records from L2: 12 MB fewer fetched, the same 0.654-0.658 ms), with 4
    // incidences per gather round for the two-plane op so that a wave (169
    // VGPRs) fits beside two one-plane waves; with 6 (197 VGPRs) it did not
    // and the step took 0.706-0.709 ms.  MPK_NEWTON_LIST2=0: two launches.
    bool list2 = true;
    if (const char* e = std::getenv("MPK_NEWTON_LIST2")) list2 = std::atoi(e) != 0;
    int side_kb = 0;
    if (const char* e = std::getenv("MPK_LIST_RING_BATCH")) {
        const int v = std::atoi(e);
        if (v == 2 || v == 3 || v == 4 || v == 6 || v == 8) side_kb = v;
    }
    if (noalloc) {
        overlap = false;
        per_cu = 16;
    }
    if (clear_flags) {
        hipLaunchKernelGGL(k_clear_tiles, dim3(grid_for(ntiles, 256)), dim3(256), 0, st, flags,
                           ntiles, flags + 2 * ntiles);
        MPK_HIP_TRY(hipGetLastError());
    }
    if (esr.plan) MPK_HIP_TRY(esr.run(0, plan->ne, st));
    // The listed passes need their list before they start: beside the class-0
    // pass, from a kernel of its own; after it, the class-0 pass's first
    // blocks build it (one launch fewer).
    if (split && overlap) {
        hipLaunchKernelGGL(k_tile_list, dim3(grid_for(ntiles, kListBlock)), dim3(kListBlock), 0,
                           st, flags, ntiles, flags + ntiles, flags + 2 * ntiles, epoch);
        MPK_HIP_TRY(hipGetLastError());
    }
    if (split) {
        const int32_t* list = flags + ntiles;
        const int32_t* count = flags + 2 * ntiles;
        // The listed tiles: K, J and R from two LDS planes (K, D), then M from
        // a pass of the sigma-mass op over the same list (one plane).  Three
        // planes in one pass (MPK_NEWTON_SPLIT_M=0) fit 6 blocks per CU; the
        // two-plane pass fits 9, the mass pass 19 (4M-tet Newton workload,
        // FP32_LOCAL: 87 us for the three-plane pass).
        auto listed = [&](hipStream_t s) {
            if (with_m && !split_m)
                return launch_ring_list(plan,
                                        FrameNewton<kPrec, true, SV>{sv, k_vals, m_vals, j_vals,
                                                                     residual, qk, qm, 0, 0},
                                        3 * sizeof(double), s, list, count, per_cu, side_kb);
            if (with_m && list2)
                return launch_ring_list2(plan,
                                         FrameNewton<kPrec, false, SV>{sv, k_vals, nullptr, j_vals,
                                                                       residual, qk, qm, 0, 0},
                                         2 * sizeof(double),
                                         FrameMass<kPrec>{sigma, m_vals, qm, 0}, sizeof(double),
                                         s, list, count, per_cu, side_kb == 6 ? 6 : 4);
            if (int rc = launch_ring_list(plan,
                                          FrameNewton<kPrec, false, SV>{sv, k_vals, nullptr,
                                                                        j_vals, residual, qk, qm,
                                                                        0, 0},
                                          2 * sizeof(double), s, list, count, per_cu, side_kb))
                return rc;
            if (!with_m) return (int)MPK_OK;
            return launch_ring_list(plan, FrameMass<kPrec>{sigma, m_vals, qm, 0}, sizeof(double),
                                    s, list, count, per_cu, side_kb);
        };
        auto class0 = [&]() {
            FrameNewtonLin<kPrec, SV> op{sv, k_vals, mfill ? nullptr : m_vals, j_vals, residual,
                                         qk, 0};
            if (!overlap) {
                op.list_flag = flags;
                op.list_out = flags + ntiles;
                op.list_count = flags + 2 * ntiles;
                op.list_ntiles = ntiles;
                op.list_epoch = epoch;
            }
            return launch_frame(plan, nullptr, op, sizeof(double), 0, st, flags, 0, epoch);
        };
        if (!overlap) {
            if (int rc = class0()) return rc;
            return listed(st);
        }
        // the listed tiles on the side stream, the class-0 pass on the
        // caller's stream; the caller's stream waits for both before anything
        // after the call
        MPK_HIP_TRY(plan->side.ensure());
        hipStream_t s2 = plan->side.s;
        MPK_HIP_TRY(hipEventRecord(plan->side.fork, st));
        MPK_HIP_TRY(hipStreamWaitEvent(s2, plan->side.fork, 0));
        const int rc = listed(s2);
        const hipError_t ej = hipEventRecord(plan->side.join, s2);
        const int rc0 = class0();
        // join after the class-0 launch, so the two passes overlap
        const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(st, plan->side.join, 0) : ej;
        if (rc) return rc;
        if (rc0) return rc0;
        MPK_HIP_TRY(ej);
        MPK_HIP_TRY(ew);
        return MPK_OK;
    }
    if (with_m)
        return launch_frame(plan, nullptr,
                            FrameNewton<kPrec, true, SV>{sv, k_vals, m_vals, j_vals, residual, qk,
                                                         qm, 0, 0},
                            3 * sizeof(double), 0, st, flags, 1);
    return launch_frame(plan, nullptr,
                        FrameNewton<kPrec, false, SV>{sv, k_vals, nullptr, j_vals, residual, qk,
                                                      qm, 0, 0},
                        2 * sizeof(double), 0, st, flags, 1);
}

// Scratch of the allocation-free Newton set (mpk_csc_newton_set_noalloc):
// the element state and the tile classes of one call.  Each device that loads
// the library holds its own copy until the process ends.  In host code the
// name is a stand-in, not a device address: only the address that
// hipGetSymbolAddress returns for the current device is passed to kernels.
constexpr size_t kNoAllocBytes = (size_t)MPK_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (on the first call there) and kept.
int scratch_address(void** out) {
    int dev = -1;
    MPK_HIP_TRY(hipGetDevice(&dev));
    MPK_REQUIRE(dev >= 0 && dev < kMaxDevices, MPK_ERR_LIMIT, "device %d beyond the %d supported",
                dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        MPK_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MPK_REQUIRE(p != nullptr, MPK_ERR_HIP, "no device address for the scratch array");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return MPK_OK;
}

}  // namespace
}  // namespace mpk

using namespace mpk;
using namespace mpk::csc;

extern "C" {

int mpk_plan_set_coords(mpk_plan* plan, const double* P, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(P || plan->nn == 0, MPK_ERR_ARGS, "NULL coordinate array");
    return load_coords(plan, P, as_stream(stream));
}

int mpk_csc_curl_curl(const mpk_plan* plan, const double* P, const double* nu,
                      const mpk_quadrature* quad, int precision, double* vals, void* stream) {
    clear_error();
    MPK_REQUIRE(plan, MPK_ERR_ARGS, "NULL plan");
    MPK_REQUIRE(plan->nnz == 0 || (nu && vals), MPK_ERR_ARGS, "NULL array argument");
    CscQuad qa;
    if (int rc = make_csc_quad(quad, true, qa)) return rc;
    if (precision != MPK_PREC_FP64_EXACT && !standard_curl_table(quad))
        precision = MPK_PREC_FP64_EXACT;
    hipStream_t st = as_stream(stream);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            return launch_exact(plan, P, ExactCurl{P, nu, vals, qa, 0}, 8, 0, st);
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
    MPK_REQUIRE(plan->nnz == 0 || vals, MPK_ERR_ARGS, "NULL array argument");
    CscQuad qa;
    if (int rc = make_csc_quad(quad, false, qa)) return rc;
    hipStream_t st = as_stream(stream);
    switch (precision) {
        case MPK_PREC_FP64_EXACT:
            return launch_exact(plan, P, ExactMass{P, coeffs, vals, qa, 0}, 8, 0, st);
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
    MPK_REQUIRE(plan->nnz == 0 || (nu && sigma && vals_c128), MPK_ERR_ARGS,
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
            return launch_exact(plan, P, ExactKM{P, nu, sigma, out, omega, qk, qm, 0}, 16, 0, st);
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
    MPK_REQUIRE(plan->n == 0 || (tags && sol_a && residual), MPK_ERR_ARGS,
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
                return launch_exact(plan, P, ExactJac<true>{P, tags, sol_a, mv, ntags, nbrk, ncoe,
                                                         j_vals, residual, qa, 0, nullptr, nullptr},
                                    8, extra, st);
            return launch_exact(plan, P, ExactJac<false>{P, tags, sol_a, mv, ntags, nbrk, ncoe,
                                                      j_vals, residual, qa, 0, nullptr, nullptr},
                                8, extra, st);
        case MPK_PREC_FP64: {
            StateView sv{};
            if (int rc = elem_state(plan, P, tags, sol_a, nullptr, mat, st, &sv)) return rc;
            if (calc_j)
                return launch_frame(plan, nullptr,
                                    FrameJac<MPK_PREC_FP64, true>{sv, j_vals, residual, qa, 0}, 8,
                                    0, st);
            return launch_frame(plan, nullptr,
                                FrameJac<MPK_PREC_FP64, false>{sv, j_vals, residual, qa, 0}, 0, 0,
                                st);
        }
        case MPK_PREC_FP32_LOCAL: {
            StateView32 sv{};
            if (int rc = elem_state(plan, P, tags, sol_a, nullptr, mat, st, &sv)) return rc;
            using J1 = FrameJac<MPK_PREC_FP32_LOCAL, true, StateView32>;
            using J0 = FrameJac<MPK_PREC_FP32_LOCAL, false, StateView32>;
            if (calc_j) return launch_frame(plan, nullptr, J1{sv, j_vals, residual, qa, 0}, 8, 0, st);
            return launch_frame(plan, nullptr, J0{sv, j_vals, residual, qa, 0}, 0, 0, st);
        }
    }
    set_error("unknown precision mode %d", precision);
    return MPK_ERR_ARGS;
}

int mpk_csc_newton_set(const mpk_plan* plan, const double* P, const int32_t* tags,
                       const double* sol_a, const mpk_material* mat, const double* sigma,
                       const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                       int precision, double* k_vals, double* m_vals, double* j_vals,
                       double* residual, void* stream) {
    clear_error();
    MPK_REQUIRE(plan && mat, MPK_ERR_ARGS, "NULL plan or material");
    MPK_REQUIRE(plan->n == 0 || (tags && sol_a), MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(precision == MPK_PREC_FP64 || precision == MPK_PREC_FP32_LOCAL, MPK_ERR_ARGS,
                "the fused Newton set has FP64 and FP32_LOCAL modes only (got %d)", precision);
    CscQuad qk, qm;
    if (int rc = make_csc_quad(quad_k, true, qk)) return rc;
    if (int rc = make_csc_quad(quad_m, false, qm)) return rc;
    MPK_REQUIRE(standard_curl_table(quad_k), MPK_ERR_ARGS,
                "the fused Newton set needs the Nedelec P1 reference curls");
    MPK_REQUIRE(mat->num_tags >= 1, MPK_ERR_ARGS, "material table without tags");
    hipStream_t st = as_stream(stream);
    if (precision == MPK_PREC_FP64)
        return newton_set<MPK_PREC_FP64, StateView>(plan, P, tags, sol_a, mat, sigma, qk, qm,
                                                    k_vals, m_vals, j_vals, residual, st);
    return newton_set<MPK_PREC_FP32_LOCAL, StateView32>(plan, P, tags, sol_a, mat, sigma, qk, qm,
                                                        k_vals, m_vals, j_vals, residual, st);
}

int mpk_csc_newton_set_noalloc(const mpk_plan* plan, const double* P, const int32_t* tags,
                               const double* sol_a, const mpk_material* mat,
                               const double* sigma, const mpk_quadrature* quad_k,
                               const mpk_quadrature* quad_m, int precision, double* k_vals,
                               double* m_vals, double* j_vals, double* residual, void* stream) {
    clear_error();
    MPK_REQUIRE(plan && mat, MPK_ERR_ARGS, "NULL plan or material");
    MPK_REQUIRE(plan->n == 0 || (tags && sol_a), MPK_ERR_ARGS, "NULL array argument");
    MPK_REQUIRE(precision == MPK_PREC_FP64 || precision == MPK_PREC_FP32_LOCAL, MPK_ERR_ARGS,
                "the fused Newton set has FP64 and FP32_LOCAL modes only (got %d)", precision);
    CscQuad qk, qm;
    if (int rc = make_csc_quad(quad_k, true, qk)) return rc;
    if (int rc = make_csc_quad(quad_m, false, qm)) return rc;
    MPK_REQUIRE(standard_curl_table(quad_k), MPK_ERR_ARGS,
                "the fused Newton set needs the Nedelec P1 reference curls");
    MPK_REQUIRE(mat->num_tags >= 1, MPK_ERR_ARGS, "material table without tags");
    // scratch layout: element state (48 B per element in FP64, 24 B in
    // FP32_LOCAL) at offset 0, then the tile classes at a 256-byte boundary
    const size_t per_elem = precision == MPK_PREC_FP64 ? 48 : 24;
    const size_t state_bytes = ((size_t)plan->ne * per_elem + 255) & ~(size_t)255;
    const int64_t ntiles = (plan->n + kTileCols - 1) / kTileCols;
    const size_t need = state_bytes + (size_t)(2 * ntiles + 1) * sizeof(int32_t);
    MPK_REQUIRE(need <= kNoAllocBytes, MPK_ERR_LIMIT,
                "the allocation-free Newton set needs %zu B of scratch for %lld elements and "
                "%lld columns; its scratch holds %zu B (use mpk_csc_newton_set)",
                need, (long long)plan->ne, (long long)plan->n, kNoAllocBytes);
    int dev = -1;
    MPK_HIP_TRY(hipGetDevice(&dev));
    MPK_REQUIRE(dev == plan->device, MPK_ERR_ARGS,
                "the plan lives on device %d but device %d is current", plan->device, dev);
    void* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    unsigned char* sb = static_cast<unsigned char*>(base);
    int32_t* flags = reinterpret_cast<int32_t*>(sb + state_bytes);
    hipStream_t st = as_stream(stream);
    if (precision == MPK_PREC_FP64)
        return newton_set<MPK_PREC_FP64, StateView>(plan, P, tags, sol_a, mat, sigma, qk, qm,
                                                    k_vals, m_vals, j_vals, residual, st, sb,
                                                    flags);
    return newton_set<MPK_PREC_FP32_LOCAL, StateView32>(plan, P, tags, sol_a, mat, sigma, qk, qm,
                                                        k_vals, m_vals, j_vals, residual, st, sb,
                                                        flags);
}

}  // extern "C"

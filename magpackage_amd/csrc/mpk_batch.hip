// Batched assembly: K + j w_f M_sigma of many meshes for many frequencies in
// one kernel launch (graph-capturable: no allocation, copy or synchronisation
// inside mpk_batch_run).
//
// Every mesh has its own plan (ring order required); the batch keeps one
// device descriptor per mesh and a tile table (mesh, local tile) over all the
// meshes' 64-column tiles.  A block takes tiles from that table -- one block
// per tile, or a persistent grid of `blocks_per_cu` blocks per CU walking
// contiguous chunks -- builds the mesh's op from its descriptor and runs the
// ring-order column pass; the finished tile (K, M) is written once per
// frequency ([f][nnz] complex128 per mesh), so the launch is output-bound.
// The schedule (ring gather batch, grid shape) is chosen by the caller's
// autotuner through mpk_batch_schedule.
#include <cstring>
#include <vector>

#include "mpk_csc_frame.hpp"
#include "mpk_plan.hpp"

namespace mpk {
namespace {

using namespace csc;

struct BatchMesh {
    TileCtx tc;
    const int2* ends;
    const double* nu;
    const double* sigma;
    double2* out;  // [nfreq][nnz]
    int64_t nnz;
};

struct BatchShared {
    CscQuad qk, qm;
    int nfreq;
    int acc_bytes_max;
};

template <int kPrec, int kB, bool kC>
__global__ void __launch_bounds__(kW)
k_csc_batch(const BatchMesh* __restrict__ meshes, const int2* __restrict__ tiles, int64_t ntiles,
            const double* __restrict__ omegas, BatchShared sh, int64_t chunk) {
    extern __shared__ __align__(16) unsigned char smem[];
    // one block per tile (chunk == 0, XCD-contiguous ranges) or contiguous
    // chunks per block of a persistent grid
    int64_t t0, t1;
    if (chunk == 0) {
        t0 = xcd_tile(blockIdx.x, gridDim.x);
        t1 = t0 + 1;
    } else {
        t0 = (int64_t)blockIdx.x * chunk;
        t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
    }
    for (int64_t t = t0; t < t1; ++t) {
        const int2 mt = tiles[t];
        const BatchMesh& bm = meshes[mt.x];
        FrameKMF<kPrec> op{FrameKM<kPrec>{bm.nu, bm.sigma, bm.out, 0.0, sh.qk, sh.qm, 0},
                           omegas, sh.nfreq, bm.nnz, sh.acc_bytes_max};
        ring_tile<FrameKMF<kPrec>, kB, kC>(bm.tc, bm.ends, op, smem, mt.y);
        __syncthreads();
    }
}

}  // namespace
}  // namespace mpk

struct mpk_batch {
    int64_t count = 0, ntiles = 0, nnz_total = 0;
    int precision = MPK_PREC_FP64;
    int ring_batch = 3;
    int blocks_per_cu = 0;
    bool compact = false;  // every plan has compact ring records
    int num_cus = 256;
    size_t lds = 0;
    mpk::BatchShared sh{};
    mpk::DevBuf meshes, tiles, omegas;
};

using namespace mpk;
using namespace mpk::csc;

extern "C" {

int mpk_batch_create(const mpk_batch_item* items, int64_t count, const double* omegas,
                     int32_t nfreq, const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                     int precision, void* stream, mpk_batch** batch_out) {
    clear_error();
    MPK_REQUIRE(batch_out && items && count >= 1 && omegas, MPK_ERR_ARGS, "bad argument");
    *batch_out = nullptr;
    MPK_REQUIRE(nfreq >= 1 && nfreq <= kMaxFreq, MPK_ERR_LIMIT, "nfreq=%d outside [1, %d]",
                nfreq, kMaxFreq);
    MPK_REQUIRE(precision == MPK_PREC_FP64 || precision == MPK_PREC_FP32_LOCAL, MPK_ERR_ARGS,
                "the batched kernel has FP64 and FP32_LOCAL modes");
    auto b = new mpk_batch();
    auto fail = [&](int rc) {
        delete b;
        return rc;
    };
    if (int rc = make_csc_quad(quad_k, true, b->sh.qk)) return fail(rc);
    if (int rc = make_csc_quad(quad_m, false, b->sh.qm)) return fail(rc);
    b->sh.nfreq = nfreq;
    b->precision = precision;
    b->count = count;
    std::vector<BatchMesh> hm((size_t)count);
    std::vector<int2> ht;
    int64_t max_tile = 0;
    for (int64_t m = 0; m < count; ++m) {
        const mpk_plan* p = items[m].plan;
        if (!p || !p->ring || !p->coords_loaded || !items[m].nu || !items[m].sigma ||
            (p->nnz > 0 && !items[m].vals_c128)) {
            set_error("batch item %lld: plan without ring order or coordinates, or NULL array",
                      (long long)m);
            return fail(MPK_ERR_ARGS);
        }
        BatchMesh& bm = hm[(size_t)m];
        if (m > 0 && p->ring8 != b->compact) {
            set_error("batch item %lld: the plans of one batch must share a ring record format",
                      (long long)m);
            return fail(MPK_ERR_ARGS);
        }
        b->compact = p->ring8;
        bm.tc = TileCtx{p->colptr,   p->inc_ptr,   p->inc_slot,   p->inc_pos,
                        p->inc_frame, p->inc_ring, p->frame_base, p->T,
                        p->P4,        p->rowind,   p->n,          p->inc_ring8,
                        p->col_head};
        bm.ends = p->col_ends;
        bm.nu = items[m].nu;
        bm.sigma = items[m].sigma;
        bm.out = reinterpret_cast<double2*>(items[m].vals_c128);
        bm.nnz = p->nnz;
        const int64_t nt = (p->n + kTileCols - 1) / kTileCols;
        for (int64_t t = 0; t < nt; ++t) ht.push_back(make_int2((int)m, (int)t));
        max_tile = p->max_tile_nnz > max_tile ? p->max_tile_nnz : max_tile;
        b->nnz_total += p->nnz;
    }
    b->ntiles = (int64_t)ht.size();
    const size_t acc_bytes = ((size_t)max_tile * 16 + 15) & ~size_t(15);
    if (acc_bytes > (size_t)kMaxLdsBytes) {
        set_error("largest tile needs %zu B of LDS (limit %d)", acc_bytes, kMaxLdsBytes);
        return fail(MPK_ERR_LIMIT);
    }
    b->lds = acc_bytes;
    b->sh.acc_bytes_max = (int)acc_bytes;
    int dev = 0, cus = 0;
    MPK_HIP_TRY(hipGetDevice(&dev));
    MPK_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    b->num_cus = cus > 0 ? cus : 256;
    hipStream_t st = as_stream(stream);
    if (b->meshes.alloc(hm.size() * sizeof(BatchMesh)) != hipSuccess ||
        b->tiles.alloc((ht.size() > 0 ? ht.size() : 1) * sizeof(int2)) != hipSuccess ||
        b->omegas.alloc((size_t)nfreq * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc failed for the batch tables");
        return fail(MPK_ERR_HIP);
    }
    hipError_t e1 = hipMemcpyAsync(b->meshes.p, hm.data(), hm.size() * sizeof(BatchMesh),
                                   hipMemcpyHostToDevice, st);
    hipError_t e2 = ht.empty() ? hipSuccess
                               : hipMemcpyAsync(b->tiles.p, ht.data(), ht.size() * sizeof(int2),
                                                hipMemcpyHostToDevice, st);
    hipError_t e3 = hipMemcpyAsync(b->omegas.p, omegas, (size_t)nfreq * sizeof(double),
                                   hipMemcpyHostToDevice, st);
    hipError_t e4 = hipStreamSynchronize(st);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
        set_error("copying the batch tables failed");
        return fail(MPK_ERR_HIP);
    }
    *batch_out = b;
    return MPK_OK;
}

int mpk_batch_schedule(mpk_batch* b, int32_t ring_batch, int32_t blocks_per_cu) {
    clear_error();
    MPK_REQUIRE(b, MPK_ERR_ARGS, "NULL batch");
    MPK_REQUIRE(ring_batch == 2 || ring_batch == 3 || ring_batch == 4 || ring_batch == 6,
                MPK_ERR_ARGS, "ring_batch %d not one of 2, 3, 4, 6", ring_batch);
    MPK_REQUIRE(blocks_per_cu >= 0 && blocks_per_cu <= 64, MPK_ERR_ARGS,
                "blocks_per_cu %d outside [0, 64]", blocks_per_cu);
    b->ring_batch = ring_batch;
    b->blocks_per_cu = blocks_per_cu;
    return MPK_OK;
}

int mpk_batch_info(const mpk_batch* b, int64_t* ntiles, int64_t* nnz_total, int64_t* lds_bytes) {
    clear_error();
    MPK_REQUIRE(b, MPK_ERR_ARGS, "NULL batch");
    if (ntiles) *ntiles = b->ntiles;
    if (nnz_total) *nnz_total = b->nnz_total;
    if (lds_bytes) *lds_bytes = (int64_t)b->lds;
    return MPK_OK;
}

int mpk_batch_run(const mpk_batch* b, void* stream) {
    clear_error();
    MPK_REQUIRE(b, MPK_ERR_ARGS, "NULL batch");
    if (b->ntiles == 0) return MPK_OK;
    hipStream_t st = as_stream(stream);
    int64_t chunk = 0;
    unsigned grid = (unsigned)b->ntiles;
    if (b->blocks_per_cu > 0) {
        const int64_t nb = (int64_t)b->num_cus * b->blocks_per_cu;
        chunk = (b->ntiles + nb - 1) / nb;
        grid = (unsigned)((b->ntiles + chunk - 1) / chunk);
    }
    const BatchMesh* ms = b->meshes.as<BatchMesh>();
    const int2* ts = b->tiles.as<int2>();
    const double* om = b->omegas.as<double>();
#define MPK_BATCH_LAUNCH(PREC, KB)                                                            \
    do {                                                                                      \
        if (b->compact)                                                                       \
            hipLaunchKernelGGL((k_csc_batch<PREC, KB, true>), dim3(grid), dim3(kW), b->lds,  \
                               st, ms, ts, b->ntiles, om, b->sh, chunk);                      \
        else                                                                                  \
            hipLaunchKernelGGL((k_csc_batch<PREC, KB, false>), dim3(grid), dim3(kW), b->lds, \
                               st, ms, ts, b->ntiles, om, b->sh, chunk);                      \
    } while (0)
    if (b->precision == MPK_PREC_FP64) {
        if (b->ring_batch == 2) MPK_BATCH_LAUNCH(MPK_PREC_FP64, 2);
        else if (b->ring_batch == 4) MPK_BATCH_LAUNCH(MPK_PREC_FP64, 4);
        else if (b->ring_batch == 6) MPK_BATCH_LAUNCH(MPK_PREC_FP64, 6);
        else MPK_BATCH_LAUNCH(MPK_PREC_FP64, 3);
    } else {
        if (b->ring_batch == 2) MPK_BATCH_LAUNCH(MPK_PREC_FP32_LOCAL, 2);
        else if (b->ring_batch == 4) MPK_BATCH_LAUNCH(MPK_PREC_FP32_LOCAL, 4);
        else if (b->ring_batch == 6) MPK_BATCH_LAUNCH(MPK_PREC_FP32_LOCAL, 6);
        else MPK_BATCH_LAUNCH(MPK_PREC_FP32_LOCAL, 3);
    }
#undef MPK_BATCH_LAUNCH
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_batch_destroy(mpk_batch* b) {
    clear_error();
    delete b;
    return MPK_OK;
}

}  // extern "C"

// Learned reluctivity predictor: a 16 -> 64 -> 64 -> 64 -> 1 tanh MLP whose
// hidden layers run on the gfx950 matrix cores.
//
// Each wave64 owns 16-row chunks.  Layer products are tiles of gfx950's BF16
// MFMA (BF16 operands, FP32 accumulation):
//   16x16x32 (v_mfma_f32_16x16x32_bf16, the 64-wide layers 2 and 3 and the
//   backward dH products): lane l holds A[l%16][8*(l/16) + i] and
//   B[8*(l/16) + i][l%16], i = 0..7;
//   16x16x16 (v_mfma_f32_16x16x16bf16_1k, the 16-wide layer 1 and the weight
//   gradients, whose reduction runs over the chunk's 16 rows): lane l holds
//   A[l%16][4*(l/16) + i], B[4*(l/16) + i][l%16], i = 0..3;
//   D (16 x 16) of both: lane l holds D[4*(l/16) + i][l%16].
// The instruction pairs A's and B's register elements by (lane group, i), so
// loading both operands with the same k numbering gives the product.
// Weights (BF16) and the chunk's activations (BF16) sit in LDS; D tiles go
// through LDS to become the next layer's A operands.
//
// Activations: tanh by a rational form (tanh_fast below), about 15 VALU
// instructions against ~40 for the library tanhf; BF16 rounding by the
// hardware conversion (round to nearest even).
//
// Training step (mpk_mlp_loss_grad): forward, MSE loss, backward through the
// three tanh layers.  Weight gradients are MFMA products with the chunk's 16
// rows as the reduction dimension, accumulated in registers across the
// wave's chunks; every wave writes its partial gradient, and a second kernel
// sums the partials in wave order (no atomics, deterministic).
//
// Rounding points (mirrored by the CPU oracle): X, hidden activations and the
// backpropagated pre-activation gradients are rounded to BF16 where they
// enter an MFMA; biases, w4, b4, tanh and all sums are FP32.
#include <cmath>

#include "mpk_common.hpp"
#include "mpk_mlp.h"

namespace mpk {
namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kIn = MPK_MLP_IN;
constexpr int kH = MPK_MLP_HIDDEN;
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kNP = MPK_MLP_NPARAM;
constexpr int kStride = 9476;  // partial-gradient stride per wave (multiple of 4)
constexpr int kMaxBlocks = 256;
// parameter offsets
constexpr int oW1 = 0, ob1 = 1024, oW2 = 1088, ob2 = 5184, oW3 = 5248, ob3 = 9344, ow4 = 9408,
              ob4 = 9472;
// BF16 copy offsets
constexpr int bW1 = 0, bW2 = 1024, bW3 = 5120;
static_assert(ob4 + 1 == kNP, "parameter layout");
static_assert(bW3 + kH * kH == MPK_MLP_NBF16, "bf16 layout");

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf4 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
    bf4 v;
    v.x = (short)a;
    v.y = (short)b;
    v.z = (short)c;
    v.w = (short)d;
    return v;
}
// 4 consecutive BF16 of a row-major LDS tile, one 8-byte read (idx is a
// multiple of 4 at every call and the tiles are 8-byte aligned)
__device__ __forceinline__ bf4 ld_row4(const uint16_t* t, int idx) {
    bf4 v;
    __builtin_memcpy(&v, t + idx, sizeof(v));
    return v;
}
// Orders the LDS accesses of one wave: the activation tiles inside the chunk
// loops belong to one wave, whose LDS instructions execute in order, so a
// fence that keeps the compiler from moving them replaces a block barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// 8 consecutive BF16 of a row-major LDS tile, one 16-byte read (idx is a
// multiple of 8 at every call and the tiles are 16-byte aligned)
__device__ __forceinline__ bf8 ld_row8(const uint16_t* t, int idx) {
    bf8 v;
    __builtin_memcpy(&v, t + idx, sizeof(v));
    return v;
}
// 4 BF16 of one column, rows r..r+3, row stride ld
__device__ __forceinline__ bf4 ld_col4(const uint16_t* t, int idx, int ld) {
    return pack4(t[idx], t[idx + ld], t[idx + 2 * ld], t[idx + 3 * ld]);
}
__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// 8 BF16 of one column, rows r..r+7, row stride ld
__device__ __forceinline__ bf8 ld_col8(const uint16_t* t, int idx, int ld) {
    uint16_t u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) u[i] = t[idx + i * ld];
    bf8 v;
    __builtin_memcpy(&v, u, sizeof(v));
    return v;
}

// tanh(x) by the rational minimax form of Eigen's generic_fast_tanh_float
// (the form XLA uses too): x P(x^2) / Q(x^2), P of degree 6 and Q of degree 3
// in x^2, on x clamped to +-7.90531110763549805.  |error| <= 2.6e-7 in exact
// arithmetic and about 4e-7 in FP32, far below the BF16 rounding (2^-9
// relative) the activation takes next.  oracle/mlp_oracle.py evaluates the
// same rational.
__device__ __forceinline__ float tanh_fast(float x) {
    constexpr float kClamp = 7.90531110763549805f;
    x = fminf(fmaxf(x, -kClamp), kClamp);
    const float x2 = x * x;
    float p = -2.76076847742355e-16f;
    p = fmaf(p, x2, 2.00018790482477e-13f);
    p = fmaf(p, x2, -8.60467152213735e-11f);
    p = fmaf(p, x2, 5.12229709037114e-08f);
    p = fmaf(p, x2, 1.48572235717979e-05f);
    p = fmaf(p, x2, 6.37261928875436e-04f);
    p = fmaf(p, x2, 4.89352455891786e-03f);
    float q = 1.19825839466702e-06f;
    q = fmaf(q, x2, 1.18534705686654e-04f);
    q = fmaf(q, x2, 2.26843463243900e-03f);
    q = fmaf(q, x2, 4.89352518554385e-03f);
    return x * p * __builtin_amdgcn_rcpf(q);
}
__device__ __forceinline__ f4 zero4() { return f4{0.f, 0.f, 0.f, 0.f}; }

struct alignas(16) Smem {
    uint16_t w1[kH * kIn];
    uint16_t w2[kH * kH];
    uint16_t w3[kH * kH];
    float b1[kH], b2[kH], b3[kH], w4[kH];
    uint16_t x[kWaves][16 * kIn];
    uint16_t h1[kWaves][16 * kH];
    uint16_t h2[kWaves][16 * kH];
    uint16_t h3[kWaves][16 * kH];
    uint16_t dz[kWaves][16 * kH];
};

__device__ void load_weights(Smem& s, const float* __restrict__ p, const uint16_t* __restrict__ wb) {
    for (int i = threadIdx.x; i < kH * kIn; i += kThreads) s.w1[i] = wb[bW1 + i];
    for (int i = threadIdx.x; i < kH * kH; i += kThreads) {
        s.w2[i] = wb[bW2 + i];
        s.w3[i] = wb[bW3 + i];
    }
    for (int i = threadIdx.x; i < kH; i += kThreads) {
        s.b1[i] = p[ob1 + i];
        s.b2[i] = p[ob2 + i];
        s.b3[i] = p[ob3 + i];
        s.w4[i] = p[ow4 + i];
    }
}

// Forward pass of one 16-row chunk (rows r0.., n valid in total).  Leaves
// BF16 activations in LDS, returns the BF16-rounded layer-3 activations of
// the lane's D entries (hb3[ob][i]: row 4g+i, unit 16 ob + c) and y of rows
// 4g+i (identical in all 16 lanes of a group).  Touches only wave w's tiles.
__device__ void forward_chunk(Smem& s, int w, const float* __restrict__ X, int64_t r0, int64_t n,
                              float b4, float (&hb3)[4][4], float (&y)[4]) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    uint16_t* xs = s.x[w];
    uint16_t* h1 = s.h1[w];
    uint16_t* h2 = s.h2[w];
    uint16_t* h3 = s.h3[w];
    for (int e = lane; e < 16 * kIn; e += 64) {
        const int r = e / kIn, k = e % kIn;
        const float v = (r0 + r < n) ? X[(r0 + r) * kIn + k] : 0.0f;
        xs[e] = f2bf(v);
    }
    wave_sync();
    // layer 1: Z1 = X W1^T + b1
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        f4 acc = mfma(ld_row4(xs, c * kIn + 4 * g), ld_row4(s.w1, (ob * 16 + c) * kIn + 4 * g),
                      zero4());
        const int o = ob * 16 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) h1[(4 * g + i) * kH + o] = f2bf(tanh_fast(acc[i] + s.b1[o]));
    }
    wave_sync();
    // layer 2
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        f4 acc = zero4();
#pragma unroll
        for (int kp = 0; kp < 2; ++kp)
            acc = mfma32(ld_row8(h1, c * kH + kp * 32 + 8 * g),
                         ld_row8(s.w2, (ob * 16 + c) * kH + kp * 32 + 8 * g), acc);
        const int o = ob * 16 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) h2[(4 * g + i) * kH + o] = f2bf(tanh_fast(acc[i] + s.b2[o]));
    }
    wave_sync();
    // layer 3 and the linear output
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        f4 acc = zero4();
#pragma unroll
        for (int kp = 0; kp < 2; ++kp)
            acc = mfma32(ld_row8(h2, c * kH + kp * 32 + 8 * g),
                         ld_row8(s.w3, (ob * 16 + c) * kH + kp * 32 + 8 * g), acc);
        const int o = ob * 16 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint16_t hb = f2bf(tanh_fast(acc[i] + s.b3[o]));
            h3[(4 * g + i) * kH + o] = hb;
            hb3[ob][i] = bf2f(hb);
            part[i] += hb3[ob][i] * s.w4[o];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = part[i];
        v += __shfl_xor(v, 8);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 1);
        y[i] = v + b4;
    }
    wave_sync();
}

__global__ void __launch_bounds__(kThreads)
k_mlp_forward(const float* __restrict__ p, const uint16_t* __restrict__ wb,
              const float* __restrict__ X, int64_t n, float* __restrict__ Y) {
    __shared__ Smem s;
    load_weights(s, p, wb);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const float b4 = p[ob4];
    const int64_t nchunk = (n + 15) / 16;
    const int64_t ngroups = (nchunk + kWaves - 1) / kWaves;
    for (int64_t cg = blockIdx.x; cg < ngroups; cg += gridDim.x) {
        const int64_t r0 = (cg * kWaves + w) * 16;
        float hb3[4][4], y[4];
        forward_chunk(s, w, X, r0, n, b4, hb3, y);
        if (c == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t r = r0 + 4 * g + i;
                if (r < n) Y[r] = y[i];
            }
        }
    }
}

__device__ __forceinline__ float group_col_sum(float v) {
    // sum over the 4 lane groups that share l % 16
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

__global__ void __launch_bounds__(kThreads)
k_mlp_grad(const float* __restrict__ p, const uint16_t* __restrict__ wb,
           const float* __restrict__ X, const float* __restrict__ T, int64_t n,
           float* __restrict__ work, float* __restrict__ loss_w) {
    __shared__ Smem s;
    load_weights(s, p, wb);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const float b4 = p[ob4];
    const float inv_n = 1.0f / (float)n;
    uint16_t* xs = s.x[w];
    uint16_t* h1 = s.h1[w];
    uint16_t* h2 = s.h2[w];
    uint16_t* dz = s.dz[w];
    f4 gW1[4], gW2[4][4], gW3[4][4];
    float gb1[4], gb2[4], gb3[4], gw4[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = zero4();
        gb1[a] = gb2[a] = gb3[a] = gw4[a] = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = gW3[a][b] = zero4();
    }
    float gb4 = 0.f, lsum = 0.f;
    const int64_t nchunk = (n + 15) / 16;
    const int64_t ngroups = (nchunk + kWaves - 1) / kWaves;
    for (int64_t cg = blockIdx.x; cg < ngroups; cg += gridDim.x) {
        const int64_t r0 = (cg * kWaves + w) * 16;
        float hb3[4][4], y[4];
        forward_chunk(s, w, X, r0, n, b4, hb3, y);
        float dy[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t r = r0 + 4 * g + i;
            const bool valid = r < n;
            const float d = valid ? y[i] - T[r] : 0.0f;
            dy[i] = 2.0f * d * inv_n;
            if (c == 0) {
                lsum += d * d;
                gb4 += dy[i];
            }
        }
        // output layer and dZ3 = dy w4 (1 - h3^2)
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            const int o = ob * 16 + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gw4[ob] += dy[i] * hb3[ob][i];
                const float z = dy[i] * s.w4[o] * (1.0f - hb3[ob][i] * hb3[ob][i]);
                gb3[ob] += z;
                dz[(4 * g + i) * kH + o] = f2bf(z);
            }
        }
        wave_sync();
        // gW3 += dZ3^T H2;  dH2 = dZ3 W3
        float dh[4][4];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                gW3[ob][kb] = mfma(ld_col4(dz, (4 * g) * kH + ob * 16 + c, kH),
                                   ld_col4(h2, (4 * g) * kH + kb * 16 + c, kH), gW3[ob][kb]);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            f4 acc = zero4();
#pragma unroll
            for (int kp = 0; kp < 2; ++kp)
                acc = mfma32(ld_row8(dz, c * kH + kp * 32 + 8 * g),
                             ld_col8(s.w3, (kp * 32 + 8 * g) * kH + kb * 16 + c, kH), acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float hb = bf2f(h2[(4 * g + i) * kH + kb * 16 + c]);
                dh[kb][i] = acc[i] * (1.0f - hb * hb);
            }
        }
        wave_sync();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gb2[kb] += dh[kb][i];
                dz[(4 * g + i) * kH + kb * 16 + c] = f2bf(dh[kb][i]);
            }
        wave_sync();
        // gW2 += dZ2^T H1;  dH1 = dZ2 W2
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                gW2[ob][kb] = mfma(ld_col4(dz, (4 * g) * kH + ob * 16 + c, kH),
                                   ld_col4(h1, (4 * g) * kH + kb * 16 + c, kH), gW2[ob][kb]);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            f4 acc = zero4();
#pragma unroll
            for (int kp = 0; kp < 2; ++kp)
                acc = mfma32(ld_row8(dz, c * kH + kp * 32 + 8 * g),
                             ld_col8(s.w2, (kp * 32 + 8 * g) * kH + kb * 16 + c, kH), acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float hb = bf2f(h1[(4 * g + i) * kH + kb * 16 + c]);
                dh[kb][i] = acc[i] * (1.0f - hb * hb);
            }
        }
        wave_sync();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gb1[kb] += dh[kb][i];
                dz[(4 * g + i) * kH + kb * 16 + c] = f2bf(dh[kb][i]);
            }
        wave_sync();
        // gW1 += dZ1^T X
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
            gW1[ob] = mfma(ld_col4(dz, (4 * g) * kH + ob * 16 + c, kH),
                           ld_col4(xs, (4 * g) * kIn + c, kIn), gW1[ob]);
        wave_sync();
    }
    // partial gradient of this wave
    float* out = work + (int64_t)(blockIdx.x * kWaves + w) * kStride;
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = ob * 16 + 4 * g + i;
            out[oW1 + o * kIn + c] = gW1[ob][i];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                out[oW2 + o * kH + kb * 16 + c] = gW2[ob][kb][i];
                out[oW3 + o * kH + kb * 16 + c] = gW3[ob][kb][i];
            }
        }
        const float s1 = group_col_sum(gb1[ob]), s2 = group_col_sum(gb2[ob]),
                    s3 = group_col_sum(gb3[ob]), s4 = group_col_sum(gw4[ob]);
        if (g == 0) {
            out[ob1 + ob * 16 + c] = s1;
            out[ob2 + ob * 16 + c] = s2;
            out[ob3 + ob * 16 + c] = s3;
            out[ow4 + ob * 16 + c] = s4;
        }
    }
    // gb4 and the loss live in the c == 0 lanes: sum over the 4 groups
    gb4 = group_col_sum(gb4);
    lsum = group_col_sum(lsum);
    if (lane == 0) {
        out[ob4] = gb4;
        loss_w[blockIdx.x * kWaves + w] = lsum;
    }
}

// grad[p] = sum over waves in order; loss = sum / n.
__global__ void __launch_bounds__(256)
k_mlp_reduce(const float* __restrict__ work, const float* __restrict__ loss_w, int nw, int64_t n,
             float* __restrict__ grad, float* __restrict__ loss) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kNP) {
        float a = 0.f;
        for (int k = 0; k < nw; ++k) a += work[(int64_t)k * kStride + i];
        grad[i] = a;
    }
    if (i == kNP) {
        float a = 0.f;
        for (int k = 0; k < nw; ++k) a += loss_w[k];
        *loss = a / (float)n;
    }
}

__device__ __forceinline__ int bf_to_param(int i) {
    return i < bW2 ? oW1 + i : (i < bW3 ? oW2 + (i - bW2) : oW3 + (i - bW3));
}

__global__ void k_mlp_sync(const float* __restrict__ p, uint16_t* __restrict__ wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < MPK_MLP_NBF16) wb[i] = f2bf(p[bf_to_param(i)]);
}

__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNP) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int grad_blocks(int64_t n) {
    const int64_t nchunk = (n + 15) / 16;
    const int64_t ngroups = (nchunk + kWaves - 1) / kWaves;
    return (int)(ngroups < kMaxBlocks ? (ngroups > 0 ? ngroups : 1) : kMaxBlocks);
}

}  // namespace
}  // namespace mpk

using namespace mpk;

extern "C" {

int mpk_mlp_sync_bf16(const float* params, uint16_t* w_bf16, void* stream) {
    clear_error();
    MPK_REQUIRE(params && w_bf16, MPK_ERR_ARGS, "NULL argument");
    hipLaunchKernelGGL(k_mlp_sync, dim3((MPK_MLP_NBF16 + 255) / 256), dim3(256), 0,
                       as_stream(stream), params, w_bf16);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_mlp_forward(const float* params, const uint16_t* w_bf16, const float* X, int64_t n,
                    float* y, void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 0, MPK_ERR_ARGS, "negative size");
    if (n == 0) return MPK_OK;
    MPK_REQUIRE(params && w_bf16 && X && y, MPK_ERR_ARGS, "NULL argument");
    const int64_t ngroups = ((n + 15) / 16 + kWaves - 1) / kWaves;
    const unsigned grid = (unsigned)(ngroups < 4096 ? ngroups : 4096);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(kThreads), 0, as_stream(stream), params,
                       w_bf16, X, n, y);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_mlp_work_size(int64_t n, int64_t* size_out) {
    clear_error();
    MPK_REQUIRE(size_out && n >= 0, MPK_ERR_ARGS, "bad argument");
    const int64_t nw = (int64_t)grad_blocks(n) * kWaves;
    *size_out = nw * kStride + nw;
    return MPK_OK;
}

int mpk_mlp_loss_grad(const float* params, const uint16_t* w_bf16, const float* X,
                      const float* t, int64_t n, float* grad, float* loss, float* work,
                      void* stream) {
    clear_error();
    MPK_REQUIRE(n >= 1, MPK_ERR_ARGS, "empty batch");
    MPK_REQUIRE(params && w_bf16 && X && t && grad && loss && work, MPK_ERR_ARGS,
                "NULL argument");
    hipStream_t st = as_stream(stream);
    const int blocks = grad_blocks(n);
    const int nw = blocks * kWaves;
    float* loss_w = work + (int64_t)nw * kStride;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kThreads), 0, st, params, w_bf16, X, t, n,
                       work, loss_w);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + 1 + 255) / 256), dim3(256), 0, st, work, loss_w,
                       nw, n, grad, loss);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

int mpk_mlp_adam(float* params, float* m, float* v, const float* grad, int64_t step, float lr,
                 float beta1, float beta2, float eps, uint16_t* w_bf16, void* stream) {
    clear_error();
    MPK_REQUIRE(params && m && v && grad && w_bf16 && step >= 1, MPK_ERR_ARGS, "bad argument");
    hipStream_t st = as_stream(stream);
    const float c1 = 1.0f - std::pow(beta1, (float)step);
    const float c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + 255) / 256), dim3(256), 0, st, params, m, v, grad,
                       lr, beta1, beta2, eps, c1, c2);
    hipLaunchKernelGGL(k_mlp_sync, dim3((MPK_MLP_NBF16 + 255) / 256), dim3(256), 0, st, params,
                       w_bf16);
    MPK_HIP_TRY(hipGetLastError());
    return MPK_OK;
}

}  // extern "C"

// The learned part of the surrogate sign mode (no counterpart in the
// reference): a 3 -> 64 -> 64 -> 64 -> 1 ReLU network that predicts a tile's
// smallest |SDF| (voxels) from its centre, trained on samples the exact
// mesh-to-volume path generates, and the tile gate that uses it.
//
// Numerics: master weights FP32 (Adam); the kernels compute with BF16 weights
// and BF16 activations (round to nearest even) and FP32 accumulation.  The
// input layer (K = 3) and the output layer (N = 1) run on the VALU, the two
// 64 x 64 layers on the matrix cores with v_mfma_f32_16x16x16_bf16.
//
// Layout: a wave processes 16 samples at a time and computes the transposed
// product Y^T = W X^T, weights as the A operand (rows = output features) and
// activations as the B operand (columns = samples).  Lane l then holds, for
// sample l & 15, the four features 16 b + 4 (l >> 4) + j (j = 0..3) of each
// 16-feature block b -- exactly the B-operand layout of the next layer's
// K-block b, so activations stay in registers from layer to layer.  The
// weight gradients sum over samples (K = samples), for which the same values
// are needed with samples along K: they go through a 16 x 64 LDS transpose.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "surrogate.hpp"

namespace bs {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NPARAM = BS_MLP_PARAMS;
// parameter offsets: W1 [64][3], b1, W2 [64][64] (row = output), b2, W3, b3, W4 [64], b4
constexpr int P_W1 = 0, P_B1 = 192, P_W2 = 256, P_B2 = 4352, P_W3 = 4416, P_B3 = 8512, P_W4 = 8576, P_B4 = 8640;
static_assert(P_B4 + 1 == NPARAM, "parameter count");

// Operand fragments of the two 64 x 64 layers: fwd[L][ob * 4 + kb][lane]
// holds W[16 ob + (lane & 15)][16 kb + 4 (lane >> 4) + j], bwd[L][ib * 4 + kb]
// [lane] holds W[16 kb + 4 (lane >> 4) + j][16 ib + (lane & 15)] (the A operand
// of W^T), j = 0..3 as four BF16 in one uint2.
struct MlpPacked {
    uint2 fwd[2][16][64];
    uint2 bwd[2][16][64];
    float w1[192];  // BF16 values
    float b1[64], b2[64], b3[64];
    float w4[64];   // BF16 values
    float b4;
};

__device__ __host__ inline uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ inline float bfround(float f) { return bf2f(f2bf(f)); }
__device__ inline uint2 pack4(const float v[4]) {
    return make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                      (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
}
__device__ inline float elem(uint2 p, int j) {
    const uint32_t w = j < 2 ? p.x : p.y;
    return bf2f((j & 1) ? (w >> 16) : (w & 0xFFFFu));
}
__device__ inline f32x4 mfma(uint2 a, uint2 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(bf16x4, a), __builtin_bit_cast(bf16x4, b), c,
                                                     0, 0, 0);
}

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ p, MlpPacked* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // 2 layers x 16 fragments x 64 lanes
    if (t < 2 * 16 * 64) {
        const int L = t >> 10, fr = (t >> 6) & 15, lane = t & 63;
        const float* W = p + (L ? P_W3 : P_W2);
        const int a = fr >> 2, kb = fr & 3, r = lane & 15, q = lane >> 4;
        float f[4], b[4];
        for (int j = 0; j < 4; ++j) {
            f[j] = W[(16 * a + r) * 64 + 16 * kb + 4 * q + j];
            b[j] = W[(16 * kb + 4 * q + j) * 64 + 16 * a + r];
        }
        out->fwd[L][fr][lane] = pack4(f);
        out->bwd[L][fr][lane] = pack4(b);
    }
    if (t < 192) out->w1[t] = bfround(p[P_W1 + t]);
    if (t < 64) {
        out->b1[t] = p[P_B1 + t];
        out->b2[t] = p[P_B2 + t];
        out->b3[t] = p[P_B3 + t];
        out->w4[t] = bfround(p[P_W4 + t]);
    }
    if (t == 0) out->b4 = p[P_B4];
}

// Activations of one 16-sample block in the lane layout described above.
struct Fwd {
    float x[3];
    uint2 h1[4], h2[4], h3[4];
    float out;
};

// One 64 x 64 layer: h_out = bf16(relu(W h_in + b)).
__device__ inline void dense64(const uint2 (*frag)[64], const float* b, const uint2 in[4], uint2 out[4], int lane) {
    const int q = lane >> 4;
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(frag[ob * 4 + kb][lane], in[kb], acc);
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = fmaxf(acc[j] + b[16 * ob + 4 * q + j], 0.0f);
        out[ob] = pack4(v);
    }
}

__device__ inline void forward(const MlpPacked& P, Fwd& F, int lane) {
    const int q = lane >> 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 16 * b + 4 * q + j;
            float z = P.b1[f];
            z = z + P.w1[3 * f] * F.x[0];
            z = z + P.w1[3 * f + 1] * F.x[1];
            z = z + P.w1[3 * f + 2] * F.x[2];
            v[j] = fmaxf(z, 0.0f);
        }
        F.h1[b] = pack4(v);
    }
    dense64(P.fwd[0], P.b2, F.h1, F.h2, lane);
    dense64(P.fwd[1], P.b3, F.h2, F.h3, lane);
    float s = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) s = s + P.w4[16 * b + 4 * q + j] * elem(F.h3[b], j);
    s = s + __shfl_xor(s, 16);
    s = s + __shfl_xor(s, 32);
    F.out = s + P.b4;
}

// Inputs: explicit points (n x 3) or the eight 4^3 tiles of each leaf
// (sample i: leaf i >> 3, tile i & 7 = x-half << 2 | y-half << 1 | z-half).
struct PointInput {
    const float* x;
    __device__ void at(uint64_t i, float p[3]) const {
        p[0] = x[3 * i];
        p[1] = x[3 * i + 1];
        p[2] = x[3 * i + 2];
    }
};
struct TileInput {
    const int32_t* origins;
    float h;
    __device__ void at(uint64_t i, float p[3]) const {
        const uint64_t leaf = i >> 3;
        const int t = (int)(i & 7);
        p[0] = ((float)(origins[4 * leaf] + ((t >> 2) & 1) * 4) + 1.5f) * h;
        p[1] = ((float)(origins[4 * leaf + 1] + ((t >> 1) & 1) * 4) + 1.5f) * h;
        p[2] = ((float)(origins[4 * leaf + 2] + (t & 1) * 4) + 1.5f) * h;
    }
};

struct InputMap {
    float sx, sy, sz, scale;
};

template <class In>
__global__ __launch_bounds__(256) void k_mlp_forward(In in, uint64_t n, const MlpPacked* __restrict__ Pg, InputMap im,
                                                     float* __restrict__ out) {
    const MlpPacked& P = *Pg;
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t n_waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t base = wave * 16; base < n; base += n_waves * 16) {
        const uint64_t i = base + (lane & 15);
        Fwd F;
        float p[3] = {0.0f, 0.0f, 0.0f};
        if (i < n) in.at(i, p);
        F.x[0] = (p[0] - im.sx) * im.scale;
        F.x[1] = (p[1] - im.sy) * im.scale;
        F.x[2] = (p[2] - im.sz) * im.scale;
        forward(P, F, lane);
        if (lane < 16 && i < n) out[i] = F.out;
    }
}

// Transposes the 16 x 64 block held in the lane layout (sample = lane & 15,
// feature 16 b + 4 (lane >> 4) + j) through LDS and returns, for feature block
// fb, the fragment with samples along K: lane l gets features 16 fb + (l & 15)
// of samples 4 (l >> 4) + j.
__device__ inline void to_lds(uint16_t* T, const uint2 v[4], int lane) {
    const int s = lane & 15, q = lane >> 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        uint16_t* row = T + s * 64 + 16 * b + 4 * q;
        row[0] = (uint16_t)(v[b].x & 0xFFFFu);
        row[1] = (uint16_t)(v[b].x >> 16);
        row[2] = (uint16_t)(v[b].y & 0xFFFFu);
        row[3] = (uint16_t)(v[b].y >> 16);
    }
}
__device__ inline uint2 from_lds(const uint16_t* T, int fb, int lane) {
    const int r = lane & 15, q = lane >> 4;
    uint32_t e[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) e[j] = T[(4 * q + j) * 64 + 16 * fb + r];
    return make_uint2(e[0] | (e[1] << 16), e[2] | (e[3] << 16));
}

// Gradient of the mean squared error over a batch.  Wave w handles blocks
// [w * bpw, (w + 1) * bpw) of 16 samples of the batch (sample k of the batch
// is perm[b0 + k], or b0 + k without perm) and writes its partial sums (FP32,
// parameter order, then the sum of squared errors) to part[w].
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ x, const float* __restrict__ y,
                                                 const uint32_t* __restrict__ perm, uint64_t b0, uint32_t nb,
                                                 float inv_batch, int bpw, const MlpPacked* __restrict__ Pg,
                                                 InputMap im, float* __restrict__ part) {
    const MlpPacked& P = *Pg;
    __shared__ uint16_t Ta[16 * 64], Tb[16 * 64];
    const int lane = threadIdx.x, s = lane & 15, q = lane >> 4;
    const int w = blockIdx.x;
    f32x4 dW2[16], dW3[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) dW2[k] = dW3[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float gW1[16][3], gb1[16], gb2[16], gb3[16], gW4[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        gW1[k][0] = gW1[k][1] = gW1[k][2] = 0.0f;
        gb1[k] = gb2[k] = gb3[k] = gW4[k] = 0.0f;
    }
    float gb4 = 0.0f, sse = 0.0f;
    for (int blk = 0; blk < bpw; ++blk) {
        const uint32_t k = (uint32_t)(w * bpw + blk) * 16u + (uint32_t)s;
        const bool valid = k < nb;
        const uint64_t i = valid ? (perm ? (uint64_t)perm[b0 + k] : b0 + k) : 0;
        Fwd F;
        F.x[0] = valid ? (x[3 * i] - im.sx) * im.scale : 0.0f;
        F.x[1] = valid ? (x[3 * i + 1] - im.sy) * im.scale : 0.0f;
        F.x[2] = valid ? (x[3 * i + 2] - im.sz) * im.scale : 0.0f;
        forward(P, F, lane);
        const float e = valid ? F.out - y[i] : 0.0f;
        const float g = 2.0f * e * inv_batch;
        if (q == 0) {
            sse += e * e;
            gb4 += g;
        }
        // output layer and the ReLU of layer 3
        uint2 dz3[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float h3 = elem(F.h3[b], j);
                gW4[4 * b + j] += g * h3;
                v[j] = h3 > 0.0f ? g * P.w4[16 * b + 4 * q + j] : 0.0f;
                gb3[4 * b + j] += v[j];
            }
            dz3[b] = pack4(v);
        }
        // dW3 += dz3^T h2 (samples along K)
        to_lds(Ta, dz3, lane);
        to_lds(Tb, F.h2, lane);
        __syncthreads();
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            const uint2 a = from_lds(Ta, ob, lane);
#pragma unroll
            for (int ib = 0; ib < 4; ++ib) dW3[ob * 4 + ib] = mfma(a, from_lds(Tb, ib, lane), dW3[ob * 4 + ib]);
        }
        __syncthreads();
        // dh2 = W3^T dz3, then the ReLU of layer 2
        uint2 dz2[4];
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma(P.bwd[1][ib * 4 + kb][lane], dz3[kb], acc);
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = elem(F.h2[ib], j) > 0.0f ? acc[j] : 0.0f;
                gb2[4 * ib + j] += v[j];
            }
            dz2[ib] = pack4(v);
        }
        // dW2 += dz2^T h1
        to_lds(Ta, dz2, lane);
        to_lds(Tb, F.h1, lane);
        __syncthreads();
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            const uint2 a = from_lds(Ta, ob, lane);
#pragma unroll
            for (int ib = 0; ib < 4; ++ib) dW2[ob * 4 + ib] = mfma(a, from_lds(Tb, ib, lane), dW2[ob * 4 + ib]);
        }
        __syncthreads();
        // dh1 = W2^T dz2, the ReLU of layer 1, then W1 and b1
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma(P.bwd[0][ib * 4 + kb][lane], dz2[kb], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float d = elem(F.h1[ib], j) > 0.0f ? acc[j] : 0.0f;
                gb1[4 * ib + j] += d;
                gW1[4 * ib + j][0] += d * F.x[0];
                gW1[4 * ib + j][1] += d * F.x[1];
                gW1[4 * ib + j][2] += d * F.x[2];
            }
        }
    }
    // per-lane sums over the wave's samples: lanes of one quarter (q) share features
    auto qsum = [](float v) {
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        return v;
    };
    float* o = part + (uint64_t)w * (NPARAM + 1);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 16 * b + 4 * q + j, k = 4 * b + j;
            const float a0 = qsum(gW1[k][0]), a1 = qsum(gW1[k][1]), a2 = qsum(gW1[k][2]);
            const float c1 = qsum(gb1[k]), c2 = qsum(gb2[k]), c3 = qsum(gb3[k]), c4 = qsum(gW4[k]);
            if (s == 0) {
                o[P_W1 + 3 * f] = a0;
                o[P_W1 + 3 * f + 1] = a1;
                o[P_W1 + 3 * f + 2] = a2;
                o[P_B1 + f] = c1;
                o[P_B2 + f] = c2;
                o[P_B3 + f] = c3;
                o[P_W4 + f] = c4;
            }
        }
    // dW blocks: D[o = 4 q + j][i = lane & 15] of block (ob, ib)
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = 16 * ob + 4 * q + j, col = 16 * ib + s;
                o[P_W2 + row * 64 + col] = dW2[ob * 4 + ib][j];
                o[P_W3 + row * 64 + col] = dW3[ob * 4 + ib][j];
            }
    float t4 = gb4, te = sse;
    for (int m = 1; m < 16; m <<= 1) {
        t4 += __shfl_xor(t4, m);
        te += __shfl_xor(te, m);
    }
    if (lane == 0) {
        o[P_B4] = t4;
        o[NPARAM] = te;
    }
}

// Sum of the waves' partials in wave order (deterministic): grad[p] and, at
// grad[NPARAM], the sum of squared errors.  With adam set, also one Adam step
// on the master weights.
struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int n_waves, float* __restrict__ grad,
                                                    int adam, AdamArgs A, float* __restrict__ params,
                                                    float* __restrict__ m, float* __restrict__ v) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float s = 0.0f;
    for (int w = 0; w < n_waves; ++w) s += part[(uint64_t)w * (NPARAM + 1) + p];
    grad[p] = s;
    if (!adam || p == NPARAM) return;
    const float mm = A.beta1 * m[p] + (1.0f - A.beta1) * s;
    const float vv = A.beta2 * v[p] + (1.0f - A.beta2) * s * s;
    m[p] = mm;
    v[p] = vv;
    params[p] -= A.lr * (mm / A.c1) / (sqrtf(vv / A.c2) + A.eps);
}

// ---------------------------------------------------------------- tile gate

__device__ inline float absf_bits(float v) { return __uint_as_float(__float_as_uint(v) & 0x7FFFFFFFu); }

// Reductions over the 16 lanes of one tile inside a wave: wave w of a leaf's
// 512 threads is the x-slab x = w, lane = y * 8 + z, so a tile's lanes share
// lane bits 5 (y half) and 2 (z half) and differ in bits 0, 1, 3 and 4.
__device__ inline unsigned tile_max(unsigned v) {
    for (int m : {1, 2, 8, 16}) v = max(v, (unsigned)__shfl_xor((int)v, m));
    return v;
}
__device__ inline unsigned tile_min(unsigned v) {
    for (int m : {1, 2, 8, 16}) v = min(v, (unsigned)__shfl_xor((int)v, m));
    return v;
}

// One workgroup of 512 threads per leaf, one thread per voxel.  Per-tile
// values are reduced within each wave first; one lane per (wave, tile)
// touches the LDS word.
__global__ __launch_bounds__(512) void k_surrogate_groups(const uint64_t* __restrict__ masks,
                                                          const float* __restrict__ values,
                                                          const float* __restrict__ pred, float thr_vox, int exact_gate,
                                                          float h, float extra, uint16_t* __restrict__ roots,
                                                          uint32_t* __restrict__ rep_cnt,
                                                          unsigned long long* __restrict__ counters) {
    __shared__ unsigned int best[8];      // per tile: largest |d| bits
    __shared__ unsigned int tmin[8];      // per tile: smallest |d| bits
    __shared__ unsigned int rep_of[8];    // per tile: the representative's offset
    __shared__ unsigned int fails[8];
    __shared__ uint32_t n_roots;
    const uint64_t leaf = blockIdx.x;
    const int v = threadIdx.x, lane = v & 63;
    const int x = v >> 6, y = (v >> 3) & 7, z = v & 7;
    const int t = ((x >> 2) << 2) | ((y >> 2) << 1) | (z >> 2);
    const bool leader = (lane & 0x1B) == 0;  // one lane per tile in the wave
    if (v < 8) {
        best[v] = 0u;
        tmin[v] = 0xFFFFFFFFu;
        rep_of[v] = 0xFFFFu;
        fails[v] = 0u;
    }
    if (v == 0) n_roots = 0;
    __syncthreads();
    const bool act = (masks[leaf * 8 + x] >> lane) & 1ull;
    const float d = act ? absf_bits(values[leaf * LEAF_VOX + v]) : 0.0f;
    const unsigned db = __float_as_uint(d);
    const unsigned wmax = tile_max(act ? db : 0u), wmin = tile_min(act ? db : 0xFFFFFFFFu);
    if (leader) {
        if (wmax) atomicMax(&best[t], wmax);
        if (wmin != 0xFFFFFFFFu) atomicMin(&tmin[t], wmin);
    }
    __syncthreads();
    // representative: the smallest offset holding the tile's largest |d|
    const unsigned mine = (act && db == best[t]) ? (unsigned)v : 0xFFFFu;
    const unsigned wrep = tile_min(mine);
    if (leader && wrep != 0xFFFFu) atomicMin(&rep_of[t], wrep);
    __syncthreads();
    const bool has = tmin[t] != 0xFFFFFFFFu;
    const bool culled = has && (exact_gate ? __uint_as_float(tmin[t]) >= thr_vox * h : pred[leaf * 8 + t] >= thr_vox);
    int root = v;
    bool failed = false;
    if (act && culled) {
        const int r = (int)rep_of[t];
        const int dx = x - (r >> 6), dy = y - ((r >> 3) & 7), dz = z - (r & 7);
        const float dist = sqrtf((float)(dx * dx + dy * dy + dz * dz)) * h;
        if (r == v || d > dist * (1.0f + 0x1p-10f) + extra) root = r;
        else failed = true;
    }
    roots[leaf * LEAF_VOX + v] = act ? (uint16_t)root : (uint16_t)0xFFFF;
    const unsigned long long fb = __ballot(failed);
    // failures of this lane's tile within the wave: its lanes are the bits of
    // the group pattern shifted to the tile's (y half, z half) corner
    const unsigned long long group = 0x0F0F0F0Full << (((lane >> 5) & 1) * 32 + ((lane >> 2) & 1) * 4);
    if (leader && (fb & group)) atomicAdd(&fails[t], (unsigned)__popcll(fb & group));
    const unsigned long long b = __ballot(act && root == v);
    if (lane == 0) atomicAdd(&n_roots, (uint32_t)__popcll(b));
    __syncthreads();
    if (v == 0) {
        rep_cnt[leaf] = n_roots;
        unsigned long long c[4] = {0, 0, 0, 0};
        for (int k = 0; k < 8; ++k) {
            if (tmin[k] == 0xFFFFFFFFu) continue;
            ++c[0];
            const bool cu = exact_gate ? __uint_as_float(tmin[k]) >= thr_vox * h : pred[leaf * 8 + k] >= thr_vox;
            if (!cu) continue;
            ++c[1];
            if (fails[k]) {
                ++c[2];
                c[3] += fails[k];
            }
        }
        for (int k = 0; k < 4; ++k)
            if (c[k]) atomicAdd(&counters[k], c[k]);
    }
}

// (tile centre, smallest |value| / h) for the eight 4^3 tiles of each leaf;
// y = -1 for a tile without an active voxel.
__global__ __launch_bounds__(512) void k_tile_samples(const int32_t* __restrict__ origins,
                                                      const uint64_t* __restrict__ masks,
                                                      const float* __restrict__ values, float h, float* __restrict__ xo,
                                                      float* __restrict__ yo) {
    __shared__ unsigned int tile_min[8];
    const uint64_t leaf = blockIdx.x;
    const int v = threadIdx.x;
    const int x = v >> 6, y = (v >> 3) & 7, z = v & 7;
    const int t = ((x >> 2) << 2) | ((y >> 2) << 1) | (z >> 2);
    if (v < 8) tile_min[v] = 0xFFFFFFFFu;
    __syncthreads();
    if ((masks[leaf * 8 + x] >> (v & 63)) & 1ull)
        atomicMin(&tile_min[t], __float_as_uint(absf_bits(values[leaf * LEAF_VOX + v])));
    __syncthreads();
    if (v < 8) {
        const uint64_t i = leaf * 8 + v;
        const TileInput in{origins, h};
        float p[3];
        in.at(i, p);
        xo[3 * i] = p[0];
        xo[3 * i + 1] = p[1];
        xo[3 * i + 2] = p[2];
        yo[i] = tile_min[v] == 0xFFFFFFFFu ? -1.0f : __uint_as_float(tile_min[v]) / h;
    }
}

InputMap input_map(const bs_mlp* m) { return InputMap{m->in_shift[0], m->in_shift[1], m->in_shift[2], m->in_scale}; }

const MlpPacked* packed(const bs_mlp* m) { return static_cast<const MlpPacked*>(m->packed); }

void repack(bs_mlp* m, hipStream_t st) {
    k_mlp_pack<<<8, 256, 0, st>>>(m->params, static_cast<MlpPacked*>(m->packed));
    BS_HIP(hipGetLastError());
}

constexpr int GRAD_BPW = 8;  // 16-sample blocks per wave in a gradient launch

// Gradient (and loss) of the batch [b0, b0 + nb) (through perm when given):
// partials then the deterministic reduction, optionally with an Adam step.
void grad_step(bs_ctx* ctx, bs_mlp* m, const float* x, const float* y, const uint32_t* perm, uint64_t b0, uint32_t nb,
               float inv_batch, float* grad_dev, const AdamArgs* adam) {
    hipStream_t st = ctx->stream;
    const int blocks = (int)div_up(nb, 16);
    const int waves = (blocks + GRAD_BPW - 1) / GRAD_BPW;
    float* part = ctx->scratch.as<float>(S_MLP_PART, (uint64_t)waves * (NPARAM + 1));
    k_mlp_grad<<<waves, 64, 0, st>>>(x, y, perm, b0, nb, inv_batch, GRAD_BPW, packed(m), input_map(m), part);
    BS_HIP(hipGetLastError());
    k_mlp_reduce<<<div_up(NPARAM + 1, 256), 256, 0, st>>>(part, waves, grad_dev, adam ? 1 : 0,
                                                           adam ? *adam : AdamArgs{}, m->params, m->adam_m,
                                                           m->adam_v);
    BS_HIP(hipGetLastError());
}

}  // namespace

void surrogate_groups(bs_ctx* ctx, hipStream_t st, const bs_volume* vol, float h, float extra, uint16_t* roots,
                      uint32_t* rep_cnt, unsigned long long* counters) {
    const uint64_t n = vol->n_leaves;
    if (!n) return;
    const bs_mlp* m = ctx->surrogate;
    float* pred = nullptr;
    if (m) {
        BS_REQUIRE(m->device == ctx->device, BS_ERR_INVALID_ARG, "the surrogate network lives on another device");
        pred = ctx->scratch.as<float>(S_SG_PRED, 8 * n);
        const uint64_t tiles = 8 * n;
        const unsigned grid = (unsigned)std::min<uint64_t>(div_up(tiles, 64), 4096);
        k_mlp_forward<TileInput><<<grid, 256, 0, st>>>(TileInput{vol->origins, h}, tiles, packed(m), input_map(m),
                                                       pred);
        BS_HIP(hipGetLastError());
    }
    k_surrogate_groups<<<(unsigned)n, 512, 0, st>>>(vol->masks, vol->values, pred, ctx->surrogate_threshold,
                                                    m ? 0 : 1, h, extra, roots, rep_cnt, counters);
    BS_HIP(hipGetLastError());
}

}  // namespace bs

using namespace bs;

namespace {
template <class F>
bs_status guard_call(F&& f) {
    try {
        f();
        set_last_error("");
        return BS_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return BS_ERR_INTERNAL;
    }
}
struct Dev {
    int prev = -1;
    explicit Dev(int d) {
        BS_HIP(hipGetDevice(&prev));
        if (prev != d) BS_HIP(hipSetDevice(d));
    }
    ~Dev() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};
}  // namespace

extern "C" {

bs_status bs_mlp_create(bs_ctx* ctx, uint64_t seed, bs_mlp** out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = nullptr;
    return guard_call([&] {
        Dev g(ctx->device);
        auto m = std::make_unique<bs_mlp>();
        m->device = ctx->device;
        BS_HIP(hipMalloc(&m->params, sizeof(float) * NPARAM));
        BS_HIP(hipMalloc(&m->adam_m, sizeof(float) * NPARAM));
        BS_HIP(hipMalloc(&m->adam_v, sizeof(float) * NPARAM));
        BS_HIP(hipMalloc(&m->packed, sizeof(MlpPacked)));
        // He-uniform weights (bound sqrt(6 / fan_in)), zero biases
        std::vector<float> p(NPARAM, 0.0f);
        std::mt19937_64 rng(seed);
        auto fill = [&](int off, int count, int fan_in) {
            std::uniform_real_distribution<float> u(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
            for (int i = 0; i < count; ++i) p[off + i] = u(rng);
        };
        fill(P_W1, 192, 3);
        fill(P_W2, 4096, 64);
        fill(P_W3, 4096, 64);
        fill(P_W4, 64, 64);
        BS_HIP(hipMemcpyAsync(m->params, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
        BS_HIP(hipMemsetAsync(m->adam_m, 0, sizeof(float) * NPARAM, ctx->stream));
        BS_HIP(hipMemsetAsync(m->adam_v, 0, sizeof(float) * NPARAM, ctx->stream));
        repack(m.get(), ctx->stream);
        BS_HIP(hipStreamSynchronize(ctx->stream));
        *out = m.release();
    });
}

void bs_mlp_destroy(bs_mlp* m) {
    if (!m) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(m->params);
    (void)hipFree(m->adam_m);
    (void)hipFree(m->adam_v);
    (void)hipFree(m->packed);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete m;
}

bs_status bs_mlp_set_params(bs_ctx* ctx, bs_mlp* m, const float* params) {
    if (!ctx || !m || !params) return BS_ERR_INVALID_ARG;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_REQUIRE(m->device == ctx->device, BS_ERR_INVALID_ARG, "network lives on another device");
        BS_HIP(hipMemcpyAsync(m->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
        BS_HIP(hipMemsetAsync(m->adam_m, 0, sizeof(float) * NPARAM, ctx->stream));
        BS_HIP(hipMemsetAsync(m->adam_v, 0, sizeof(float) * NPARAM, ctx->stream));
        m->adam_t = 0;
        repack(m, ctx->stream);
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_mlp_get_params(bs_ctx* ctx, const bs_mlp* m, float* params) {
    if (!ctx || !m || !params) return BS_ERR_INVALID_ARG;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_HIP(hipMemcpyAsync(params, m->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_mlp_set_input_map(bs_mlp* m, float sx, float sy, float sz, float scale) {
    if (!m || !std::isfinite(scale) || !std::isfinite(sx) || !std::isfinite(sy) || !std::isfinite(sz))
        return BS_ERR_INVALID_ARG;
    m->in_shift[0] = sx;
    m->in_shift[1] = sy;
    m->in_shift[2] = sz;
    m->in_scale = scale;
    return BS_OK;
}

bs_status bs_mlp_forward_device(bs_ctx* ctx, const bs_mlp* m, const float* x_dev, size_t n, float* out_dev) {
    if (!ctx || !m || (n && (!x_dev || !out_dev))) return BS_ERR_INVALID_ARG;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_REQUIRE(m->device == ctx->device, BS_ERR_INVALID_ARG, "network lives on another device");
        if (!n) return;
        const unsigned grid = (unsigned)std::min<uint64_t>(div_up(n, 64), 4096);
        k_mlp_forward<PointInput><<<grid, 256, 0, ctx->stream>>>(PointInput{x_dev}, n, packed(m), input_map(m),
                                                                 out_dev);
        BS_HIP(hipGetLastError());
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_mlp_gradient_device(bs_ctx* ctx, bs_mlp* m, const float* x_dev, const float* y_dev, size_t n,
                                 float* grad, double* loss) {
    if (!ctx || !m || !x_dev || !y_dev || !grad || n == 0) return BS_ERR_INVALID_ARG;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_REQUIRE(m->device == ctx->device, BS_ERR_INVALID_ARG, "network lives on another device");
        BS_REQUIRE(n < (1ull << 31), BS_ERR_INVALID_ARG, "at most 2^31 samples");
        float* gd = ctx->scratch.as<float>(S_MLP_GRAD, NPARAM + 1);
        grad_step(ctx, m, x_dev, y_dev, nullptr, 0, (uint32_t)n, 1.0f / (float)n, gd, nullptr);
        std::vector<float> h(NPARAM + 1);
        BS_HIP(hipMemcpyAsync(h.data(), gd, sizeof(float) * (NPARAM + 1), hipMemcpyDeviceToHost, ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
        std::copy(h.begin(), h.begin() + NPARAM, grad);
        if (loss) *loss = (double)h[NPARAM] / (double)n;
    });
}

bs_status bs_mlp_train_device(bs_ctx* ctx, bs_mlp* m, const float* x_dev, const float* y_dev, size_t n,
                              const bs_mlp_train_params* tp, float* epoch_loss) {
    if (!ctx || !m || !x_dev || !y_dev || !tp || n == 0) return BS_ERR_INVALID_ARG;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_REQUIRE(m->device == ctx->device, BS_ERR_INVALID_ARG, "network lives on another device");
        BS_REQUIRE(n < (1ull << 31), BS_ERR_INVALID_ARG, "at most 2^31 samples");
        BS_REQUIRE(tp->batch > 0 && tp->epochs > 0 && tp->lr > 0.0f, BS_ERR_INVALID_ARG, "bad training parameters");
        hipStream_t st = ctx->stream;
        uint32_t* perm = ctx->scratch.as<uint32_t>(S_MLP_PERM, n);
        float* gd = ctx->scratch.as<float>(S_MLP_GRAD, NPARAM + 1);
        std::vector<uint32_t> order(n);
        std::mt19937_64 rng(tp->seed);
        const uint64_t steps_per_epoch = (n + tp->batch - 1) / tp->batch;
        const uint64_t total = steps_per_epoch * tp->epochs;
        std::vector<float> hl(1);
        uint64_t step = 0;
        for (uint32_t ep = 0; ep < tp->epochs; ++ep) {
            std::iota(order.begin(), order.end(), 0u);
            std::shuffle(order.begin(), order.end(), rng);
            BS_HIP(hipMemcpyAsync(perm, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
            double sse = 0.0;
            for (uint64_t b0 = 0; b0 < n; b0 += tp->batch, ++step) {
                const uint32_t nb = (uint32_t)std::min<uint64_t>(tp->batch, n - b0);
                // cosine decay from lr to lr_final
                const double f = 0.5 * (1.0 + std::cos(3.14159265358979323846 * (double)step / (double)total));
                AdamArgs A;
                A.lr = (float)(tp->lr_final + (tp->lr - tp->lr_final) * f);
                A.beta1 = tp->beta1;
                A.beta2 = tp->beta2;
                A.eps = tp->eps;
                ++m->adam_t;
                A.c1 = (float)(1.0 - std::pow((double)tp->beta1, (double)m->adam_t));
                A.c2 = (float)(1.0 - std::pow((double)tp->beta2, (double)m->adam_t));
                grad_step(ctx, m, x_dev, y_dev, perm, b0, nb, 1.0f / (float)nb, gd, &A);
                repack(m, st);
                if (epoch_loss) {
                    BS_HIP(hipMemcpyAsync(hl.data(), gd + NPARAM, sizeof(float), hipMemcpyDeviceToHost, st));
                    BS_HIP(hipStreamSynchronize(st));
                    sse += hl[0];
                }
            }
            if (epoch_loss) epoch_loss[ep] = (float)(sse / (double)n);
        }
        BS_HIP(hipStreamSynchronize(st));
    });
}

bs_status bs_volume_tile_samples(bs_ctx* ctx, const bs_volume* v, float* x_dev, float* y_dev, size_t capacity,
                                 size_t* n_out) {
    if (!ctx || !v || !n_out) return BS_ERR_INVALID_ARG;
    *n_out = 0;
    return guard_call([&] {
        Dev g(ctx->device);
        BS_REQUIRE(v->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
        const uint64_t n = 8 * v->n_leaves;
        *n_out = n;
        BS_REQUIRE(n <= capacity, BS_ERR_CAPACITY, "tile sample buffers hold fewer than 8 entries per leaf");
        if (!n) return;
        BS_REQUIRE(x_dev && y_dev, BS_ERR_INVALID_ARG, "null sample buffer");
        k_tile_samples<<<(unsigned)v->n_leaves, 512, 0, ctx->stream>>>(v->origins, v->masks, v->values,
                                                                       v->voxel_size, x_dev, y_dev);
        BS_HIP(hipGetLastError());
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_ctx_set_surrogate(bs_ctx* ctx, const bs_mlp* m, float threshold_vox) {
    if (!ctx || !std::isfinite(threshold_vox)) return BS_ERR_INVALID_ARG;
    ctx->surrogate = m;
    ctx->surrogate_threshold = threshold_vox;
    return BS_OK;
}

}  // extern "C"

// Learned proposal for the tracklet stage: a 24-64-64-64-1 MLP (ReLU hidden
// layers, sigmoid output) that scores (inner SG, outer-inner MD) pairs that
// passed the first-stage tracklet cuts.  This component has no counterpart in
// the reference; the exact cuts still decide every object, the network only
// drops pairs.
//
//  * inference: BF16 MFMA (v_mfma_f32_16x16x32_bf16, gfx950) with FP32
//    accumulation on 64-row tiles of BF16 feature rows in LDS.  Layers are
//    computed transposed, D = W . H^T, 16 neurons x 16 samples per product.
//    Lane l = 16 g + c holds accumulator rows 4g..4g+3 of sample c; the B
//    operand of a 32-deep K step holds 8 K values per lane (slots 8g..8g+7).
//    The hidden units are assigned to K slots so that the accumulators of
//    output tiles 2u and 2u+1 are, after ReLU and rounding, exactly the lane's
//    B operand of the next layer's K step u: activations never leave registers
//    (the host packs the weights in that order).
//  * training: FP32 forward and backward kernels (weighted binary cross
//    entropy on logits, gradients reduced per block in LDS, Adam update).
#pragma once

#include <hip/hip_runtime.h>

namespace sdl {
namespace mlp {

constexpr int kIn = 24;    // features
constexpr int kH = 64;     // hidden width
constexpr int kTiles = kH / 16;
constexpr int kSteps = kH / 32;  // K steps of a hidden layer
// parameter layout (float32, weights as [out][in])
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kH * kIn;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH * kH;
constexpr int kW4 = kB3 + kH;
constexpr int kB4 = kW4 + kH;
constexpr int kNParams = kB4 + 1;  // 9985

typedef short short4v __attribute__((ext_vector_type(4)));
typedef short short8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

// The hidden unit in K slot (step u, lane group g, element i) of layers 2 and 3.
__host__ __device__ constexpr int hiddenOfSlot(int u, int g, int i) { return 16 * (2 * u + i / 4) + 4 * g + (i % 4); }

// Weight fragments in the lane order of the MFMA operands (packed on the host),
// and the vectors that every lane group reads at its own accumulator rows.
struct Packed {
    short8v a1[kTiles][64];          // layer 1: [output tile][lane], K = 32
    short8v a2[kTiles][kSteps][64];  // layer 2: [output tile][K step][lane]
    short8v a3[kTiles][kSteps][64];
    float b2[kH], b3[kH], w4[kH];
    float b4;
};

__device__ __forceinline__ short bf16Bits(float f) {
    const __bf16 h = (__bf16)f;  // round to nearest even
    return __builtin_bit_cast(short, h);
}

__device__ __forceinline__ short4v reluBf16(float4v d) {
    short4v r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = bf16Bits(fmaxf(d[j], 0.f));
    return r;
}

__device__ __forceinline__ bf16x8 join(short4v lo, short4v hi) {
    const short8v v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ float4v mfma(short8v a, bf16x8 b, float4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), b, c, 0, 0, 0);
}

// Per-wave copy of the layer-2 weight fragments (one lane's share of the A
// operands); layers 1 and 3 are read from the block's LDS copy (Vecs) at each
// use, which keeps the register count low enough for four waves per SIMD.
struct Frags {
    short8v a2[kTiles][kSteps];
    float b4;
};

__device__ __forceinline__ void loadFrags(const Packed* P, int lane, Frags& F) {
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int u = 0; u < kSteps; ++u) F.a2[t][u] = P->a2[t][u][lane];
    F.b4 = P->b4;
}

// Layer-1 and layer-3 fragments, biases of layers 2 and 3 and the output
// weights in LDS (one copy per block).  All threads must call.
struct Vecs {
    short8v a1[kTiles][64];
    short8v a3[kTiles][kSteps][64];
    float b2[kH], b3[kH], w4[kH];
};

__device__ __forceinline__ void loadVecs(const Packed* P, Vecs& V) {
    for (int k = threadIdx.x; k < kTiles * 64; k += blockDim.x) {
        const int t = k / 64, l = k % 64;
        V.a1[t][l] = P->a1[t][l];
        for (int u = 0; u < kSteps; ++u) V.a3[t][u][l] = P->a3[t][u][l];
    }
    for (int k = threadIdx.x; k < kH; k += blockDim.x) {
        V.b2[k] = P->b2[k];
        V.b3[k] = P->b3[k];
        V.w4[k] = P->w4[k];
    }
    __syncthreads();
}

__device__ __forceinline__ float4v at4(const float* v, int i) { return *reinterpret_cast<const float4v*>(v + i); }

__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A feature row: 32 BF16 (24 features, the constant 1 at 24 carries b1, zeros).
__device__ __forceinline__ void featureRow(const float* f, short8v (&v)[4]) {
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k / 8][k % 8] = bf16Bits(k < kIn ? f[k] : (k == kIn ? 1.f : 0.f));
}

// Logits of 16 rows: x is this lane's B fragment (elements 8g..8g+7 of row c,
// lane = 16 g + c); every lane returns the logit of row c.  All 64 lanes must
// call.  A row only affects its own logit (every MFMA column is one row), so
// rows past the end may hold anything.
__device__ __forceinline__ float logit16(bf16x8 x, int lane, const Frags& F, const Vecs& V) {
    // an opaque copy of the lane index: the LDS operands (fragments, biases,
    // output weights) are read at each use instead of being hoisted into
    // registers by the compiler
    int l = lane;
    asm volatile("" : "+v"(l));
    const int g = l >> 4;
    short4v h1[kTiles], h2[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) h1[t] = reluBf16(mfma(V.a1[t][l], x, float4v{0.f, 0.f, 0.f, 0.f}));
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        float4v d = at4(V.b2, 16 * t + 4 * g);
#pragma unroll
        for (int u = 0; u < kSteps; ++u) d = mfma(F.a2[t][u], join(h1[2 * u], h1[2 * u + 1]), d);
        h2[t] = reluBf16(d);
    }
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        float4v d = at4(V.b3, 16 * t + 4 * g);
#pragma unroll
        for (int u = 0; u < kSteps; ++u) d = mfma(V.a3[t][u][l], join(h2[2 * u], h2[2 * u + 1]), d);
        const float4v w = at4(V.w4, 16 * t + 4 * g);
#pragma unroll
        for (int v = 0; v < 4; ++v) part += w[v] * fmaxf(d[v], 0.f);
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + F.b4;
}

// Logits of 64 rows: x[st] is this lane's B fragment of rows st * 16 + c.
// Lane l returns row l's logit.  All 64 lanes must call.
__device__ __forceinline__ float logit64(const bf16x8 (&x)[4], int lane, const Frags& F, const Vecs& V) {
    float mine = 0.f;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const float z = logit16(x[st], lane, F, V);
        if ((lane >> 4) == st) mine = z;  // lane st * 16 + c holds row st * 16 + c
    }
    return mine;
}

// Logits of n rows of 4 short8v each (row-major, 64 B per row), one wave per 64
// rows; the row buffer holds at least a whole number of 64-row tiles.
__device__ __forceinline__ void loadTile(const short8v* rows, long long tile, int lane, bf16x8 (&x)[4]) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int st = 0; st < 4; ++st) x[st] = __builtin_bit_cast(bf16x8, rows[(tile * 64 + st * 16 + c) * 4 + g]);
}

// Inference on stored features (row-major n x 24, FP32): one wave per 64 rows;
// the rows are rounded to BF16 into the wave's LDS tile.
__global__ void __launch_bounds__(256) k_mlp_forward(const float* X, long long n, const Packed* P, float* logit) {
    __shared__ __align__(16) short8v tiles[4][64 * 5];  // 64 rows of 4 short8v, padded to 5
    __shared__ __align__(16) Vecs V;
    loadVecs(P, V);
    const int lane = threadIdx.x & 63;
    short8v* tile = tiles[threadIdx.x >> 6];
    Frags F;
    loadFrags(P, lane, F);
    const long long nTiles = (n + 63) / 64;
    const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int g = lane >> 4, c = lane & 15;
    for (long long t = w0; t < nTiles; t += nw) {
        const long long i = t * 64 + lane;
        float f[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) f[k] = i < n ? X[i * kIn + k] : 0.f;
        short8v v[4];
        featureRow(f, v);
#pragma unroll
        for (int q = 0; q < 4; ++q) tile[lane * 5 + q] = v[q];
        waveSync();
        bf16x8 x[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) x[st] = __builtin_bit_cast(bf16x8, tile[(st * 16 + c) * 5 + g]);
        waveSync();  // the tile is rewritten by the next iteration
        const float z = logit64(x, lane, F, V);
        if (i < n) logit[i] = z;
    }
}

// ---------------------------------------------------------------- training
constexpr int kTrainSamples = 32;   // samples per block
constexpr int kTrainThreads = 256;  // 8 threads per sample
constexpr int kPerSample = kTrainThreads / kTrainSamples;

// One block: forward and backward of kTrainSamples samples (rows idx[base + s],
// or base + s when idx is null) of the mean weighted BCE over `nTotal` samples;
// gradients and loss are added to grad[] / loss[0].
__global__ void __launch_bounds__(kTrainThreads) k_mlp_fwd_bwd(const float* X, const float* Y, const int* idx,
                                                               long long base, long long nBatch, long long nTotal,
                                                               const float* params, float posWeight, float* grad,
                                                               float* loss) {
    __shared__ float W[kNParams];
    __shared__ float x[kTrainSamples][kIn];
    __shared__ float h1[kTrainSamples][kH], h2[kTrainSamples][kH], h3[kTrainSamples][kH];
    __shared__ float d1[kTrainSamples][kH], d2[kTrainSamples][kH], d3[kTrainSamples][kH];
    __shared__ float d4[kTrainSamples];
    const int tid = threadIdx.x;
    for (int k = tid; k < kNParams; k += kTrainThreads) W[k] = params[k];
    const int s = tid / kPerSample, r = tid % kPerSample;  // sample, part
    const long long row = (long long)blockIdx.x * kTrainSamples + s;
    const bool valid = row < nBatch;
    const long long src = valid ? (idx ? (long long)idx[base + row] : base + row) : 0;
    for (int k = r; k < kIn; k += kPerSample) x[s][k] = valid ? X[src * kIn + k] : 0.f;
    __syncthreads();
    // forward: each part computes kH / kPerSample neurons of its sample
    constexpr int kNpp = kH / kPerSample;
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) {
        float a = W[kB1 + j];
        for (int k = 0; k < kIn; ++k) a += W[kW1 + j * kIn + k] * x[s][k];
        h1[s][j] = fmaxf(a, 0.f);
    }
    __syncthreads();
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) {
        float a = W[kB2 + j];
        for (int k = 0; k < kH; ++k) a += W[kW2 + j * kH + k] * h1[s][k];
        h2[s][j] = fmaxf(a, 0.f);
    }
    __syncthreads();
    float zp = 0.f;
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) {
        float a = W[kB3 + j];
        for (int k = 0; k < kH; ++k) a += W[kW3 + j * kH + k] * h2[s][k];
        h3[s][j] = fmaxf(a, 0.f);
        zp += W[kW4 + j] * h3[s][j];
    }
    // the kPerSample parts of a sample are consecutive lanes
    for (int o = 1; o < kPerSample; o <<= 1) zp += __shfl_xor(zp, o);
    if (r == 0) {
        const float z = zp + W[kB4];
        const float y = valid ? Y[src] : 0.f;
        const float w = valid ? (y > 0.5f ? posWeight : 1.f) / float(nTotal) : 0.f;
        const float p = 1.f / (1.f + expf(-z));
        d4[s] = w * (p - y);
        const float l = w * (fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z))));
        if (valid) atomicAdd(loss, l);
    }
    __syncthreads();
    // backward through the hidden layers (ReLU masks from the activations)
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) d3[s][j] = h3[s][j] > 0.f ? d4[s] * W[kW4 + j] : 0.f;
    __syncthreads();
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) {
        float a = 0.f;
        for (int i = 0; i < kH; ++i) a += W[kW3 + i * kH + j] * d3[s][i];
        d2[s][j] = h2[s][j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int j = r * kNpp; j < (r + 1) * kNpp; ++j) {
        float a = 0.f;
        for (int i = 0; i < kH; ++i) a += W[kW2 + i * kH + j] * d2[s][i];
        d1[s][j] = h1[s][j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // parameter gradients: sums over the block's samples, one entry per thread at a time
    for (int e = tid; e < kNParams; e += kTrainThreads) {
        float g = 0.f;
        if (e < kB1) {
            const int j = e / kIn, k = e % kIn;
            for (int q = 0; q < kTrainSamples; ++q) g += d1[q][j] * x[q][k];
        } else if (e < kW2) {
            for (int q = 0; q < kTrainSamples; ++q) g += d1[q][e - kB1];
        } else if (e < kB2) {
            const int j = (e - kW2) / kH, k = (e - kW2) % kH;
            for (int q = 0; q < kTrainSamples; ++q) g += d2[q][j] * h1[q][k];
        } else if (e < kW3) {
            for (int q = 0; q < kTrainSamples; ++q) g += d2[q][e - kB2];
        } else if (e < kB3) {
            const int j = (e - kW3) / kH, k = (e - kW3) % kH;
            for (int q = 0; q < kTrainSamples; ++q) g += d3[q][j] * h2[q][k];
        } else if (e < kW4) {
            for (int q = 0; q < kTrainSamples; ++q) g += d3[q][e - kB3];
        } else if (e < kB4) {
            for (int q = 0; q < kTrainSamples; ++q) g += d4[q] * h3[q][e - kW4];
        } else {
            for (int q = 0; q < kTrainSamples; ++q) g += d4[q];
        }
        if (g != 0.f) atomicAdd(&grad[e], g);
    }
}

// Adam step on every parameter; clears the gradient.
__global__ void k_mlp_adam(float* params, float* grad, float* m, float* v, float lr, float beta1, float beta2,
                           float eps, float bc1, float bc2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kNParams) return;
    const float g = grad[e];
    const float mm = beta1 * m[e] + (1.f - beta1) * g;
    const float vv = beta2 * v[e] + (1.f - beta2) * g * g;
    m[e] = mm;
    v[e] = vv;
    params[e] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    grad[e] = 0.f;
}

}  // namespace mlp
}  // namespace sdl

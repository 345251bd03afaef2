// mlp.h -- the warm-start network (16 -> 64 -> 64 -> 1, SiLU) on bf16 MFMA with FP32 accumulation.
#pragma once

#include "rx_common.h"

constexpr int kMlpIn = 16;   // 12 features + 4 zero padding (one MFMA K step)
constexpr int kMlpH = 64;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Device view of the network used by the kernels.  Weights are stored [out][in] (row = output unit), which is
// the B-fragment order of v_mfma_f32_16x16x16_bf16 (4 consecutive inputs of one output per lane).
struct MlpDev {
    const unsigned short* w1;   // bf16 [64][16]
    const unsigned short* w2;   // bf16 [64][64]
    const unsigned short* w2t;  // bf16 [64][64], transposed (backward)
    const float* b1;            // [64]
    const float* b2;            // [64]
    const float* w3;            // [64]
    const float* b3;            // [1]
    const float* in_mu;         // [16] feature standardisation
    const float* in_isd;        // [16]
    float out_scale, out_bias;  // target = y * out_scale + out_bias
};

// ------------------------------------------------------------------------------------------ device helpers
namespace mlpd {
// SiLU with the hardware exp2 and reciprocal (one v_exp_f32, one v_rcp_f32): the activations are rounded to
// bf16 for the next layer, far coarser than the few-ulp error of the fast division.
__device__ inline float silu(float z) { return __fdividef(z, 1.0f + __expf(-z)); }
__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ inline bf16x4 ld_bf4(const unsigned short* p) {  // 4 consecutive bf16 (8-byte aligned)
    uint2 u = *(const uint2*)p;
    bf16x4 r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}
// Four floats to bf16 with the hardware's round-to-nearest-even conversion (two v_cvt_pk_bf16_f32).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
__device__ inline bf16x4 pack4_hw(float a, float b, float c, float d) {
    const f32x2_t lo = {a, b}, hi = {c, d};
    const u32x2_t u = {__builtin_bit_cast(unsigned, __builtin_convertvector(lo, bf16x2_t)),
                       __builtin_bit_cast(unsigned, __builtin_convertvector(hi, bf16x2_t))};
    return __builtin_bit_cast(bf16x4, u);
}
}  // namespace mlpd

constexpr int kMlpW2Stride = kMlpH + 8;  // padded LDS row: the 16 rows a fragment read touches fall in different banks

// The network staged in LDS by one block for inference inside another kernel.
// Bound on the learned first multiplier: with kappa > 0 it stays between 0 and kappa times the plain first
// multiplier lam0 (the same sign, at most kappa times its size); kappa = 0 leaves it as predicted.
__device__ inline float warm_clamp(float lam, float lam0, float kappa) {
    if (!(kappa > 0.0f)) return lam;
    const float b = kappa * lam0;
    return fminf(fmaxf(lam, fminf(b, 0.0f)), fmaxf(b, 0.0f));
}

struct MlpShared {
    __align__(16) unsigned short w2[kMlpH * kMlpW2Stride];  // bf16 [unit][input]
    __align__(16) float sb[3 * kMlpH];                       // b1, b2, w3
    __align__(16) unsigned short w1[kMlpH * kMlpIn];         // bf16 [unit][input]
    __align__(16) float mu[kMlpIn];
    __align__(16) float isd[kMlpIn];
};

// Copies the network into s; every thread of the block calls it, and a __syncthreads must follow.
__device__ inline void mlp_stage(const MlpDev& m, MlpShared& s) {
    for (int t = threadIdx.x; t < 3 * kMlpH; t += blockDim.x)
        s.sb[t] = t < kMlpH ? m.b1[t] : (t < 2 * kMlpH ? m.b2[t - kMlpH] : m.w3[t - 2 * kMlpH]);
    for (int t = threadIdx.x; t < kMlpH * kMlpH / 4; t += blockDim.x) {
        const int row = t / (kMlpH / 4), c4 = t % (kMlpH / 4);
        ((uint2*)(s.w2 + row * kMlpW2Stride))[c4] = ((const uint2*)m.w2)[t];
    }
    for (int t = threadIdx.x; t < kMlpH * kMlpIn / 4; t += blockDim.x) ((uint2*)s.w1)[t] = ((const uint2*)m.w1)[t];
    for (int t = threadIdx.x; t < kMlpIn; t += blockDim.x) {
        s.mu[t] = m.in_mu[t];
        s.isd[t] = m.in_isd[t];
    }
}

// The lane's 12 features (f[0..11]) standardised and stored as 16 bf16 in its LDS row (features 12..15 are the
// zero padding), as k_mlp_infer forms its input fragments.
__device__ inline void mlp_write_row(const float* f, unsigned short* xrow, const MlpShared& s) {
    using namespace mlpd;
    // the staged constants are read here, not hoisted into the caller's loops (where they would hold registers)
    asm volatile("" ::: "memory");
    uint2* dst = (uint2*)xrow;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float x[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 4 * q + e;
            x[e] = ((c < 12 ? f[c] : 0.0f) - s.mu[c]) * s.isd[c];
        }
        dst[q] = __builtin_bit_cast(uint2, pack4_hw(x[0], x[1], x[2], x[3]));
    }
}

// Inference for the 64 particles of a wavefront, one per lane, from their rows (xrows: this wave's 64 x 16 bf16
// LDS rows written by mlp_write_row), with the same operations as k_mlp_infer: four 16-row tiles in transposed form
// (H^T = W X^T, weights as the A operand).  Every lane of the wave must call it; returns the lane's network output
// y * out_scale + out_bias.
__device__ inline float mlp_predict_rows(const unsigned short* xrows, const MlpShared& s, const MlpDev& m) {
    using namespace mlpd;
    const int l = threadIdx.x & 63, i = l & 15, kb = l >> 4;
    __builtin_amdgcn_s_waitcnt(0);  // the rows are read by other lanes of the wave: let the LDS writes land first
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");  // no weight read moves above this point
    bf16x4 w1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w1[t] = ld_bf4(s.w1 + (16 * t + i) * kMlpIn + 4 * kb);  // A = W1[16t + i][4kb + e]
    const float4* sb4 = (const float4*)s.sb;  // sb4[4t + kb] = units 16t + 4kb .. + 3
    float mine = 0.0f;
#pragma unroll 1
    for (int tile = 0; tile < 4; ++tile) {  // rolled: keeps the caller's register budget
        const bf16x4 xb = ld_bf4(xrows + (16 * tile + i) * kMlpIn + 4 * kb);  // B = X^T[4kb + e][row i]
        bf16x4 h1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 d = {0, 0, 0, 0};
            d = mfma(w1[t], xb, d);  // D[unit 4kb + e][row i] of units 16t..16t+15
            const float4 b = sb4[4 * t + kb];
            h1[t] = pack4_hw(silu(d[0] + b.x), silu(d[1] + b.y), silu(d[2] + b.z), silu(d[3] + b.w));
        }
        float p = 0.0f;
#pragma unroll 1
        for (int t = 0; t < 4; ++t) {  // rolled: one unit tile's W2 fragments and biases in registers at a time
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)  // A = W2[16t + i][16ks + 4kb + e]
                acc = mfma(*(const bf16x4*)(s.w2 + (16 * t + i) * kMlpW2Stride + 16 * ks + 4 * kb), h1[ks], acc);
            const float4 b = sb4[16 + 4 * t + kb], w = sb4[32 + 4 * t + kb];
            p += silu(acc[0] + b.x) * w.x + silu(acc[1] + b.y) * w.y + silu(acc[2] + b.z) * w.z + silu(acc[3] + b.w) * w.w;
        }
        p += __shfl_xor(p, 16, 64);
        p += __shfl_xor(p, 32, 64);
        if (kb == tile) mine = p;  // lane 16 tile + i is particle i of this tile
    }
    return (mine + m.b3[0]) * m.out_scale + m.out_bias;
}

// Inference over n feature rows (16 floats each): xl[i].w = (y_i * out_scale + out_bias) * w_scale.
void rx_mlp_infer_to_w(const float* feat, float4* xl, uint n, const MlpDev& m, float w_scale, float kappa, hipStream_t s);
// Inference into a plain array: y[i] = y_i * out_scale + out_bias.
void rx_mlp_infer(const float* feat, float* y, uint n, const MlpDev& m, hipStream_t s);
// Copies the device view out of an rxMLP object (returns false for a null object).  The view borrows the
// object's memory: it is valid while the object lives and reflects later SetParams / SetNormalization calls.
bool rx_mlp_view(void* mlp_object, MlpDev* out);
// Bytes of an inference copy of a network (weights, biases and standardisation; no training state).
size_t rx_mlp_copy_bytes();
// Copies the inference weights of an rxMLP object into dst (device memory of rx_mlp_copy_bytes() bytes) on
// stream s, after the object's pending work has finished, and points *out at the copy: the copy stays valid
// after the object is changed or destroyed.
bool rx_mlp_copy(void* mlp_object, void* dst, hipStream_t s, MlpDev* out);

// mlp.hip -- forward, backward and Adam kernels of the warm-start network on bf16 MFMA (FP32 accumulation).
//
// Network: x (16, standardised) -> 64 -> 64 -> 1 with SiLU, y standardised target.
// Matrix products use v_mfma_f32_16x16x16_bf16 (one wavefront computes a 16 x 16 tile, K = 16 per
// instruction).  Fragment layouts of that instruction (wave64):
//   A (16 x 16, rows i, cols k):  lane l holds A[l % 16][4 (l / 16) + e], e = 0..3
//   B (16 x 16, rows k, cols j):  lane l holds B[4 (l / 16) + e][l % 16]
//   D (16 x 16):                  lane l holds D[4 (l / 16) + e][l % 16]
// Activations change from the D layout to the A layout through LDS.
#include <math.h>
#include <string.h>

#include <random>
#include <vector>

#include "mlp.h"
#include "../../include/rx_mlp.h"

namespace {

constexpr int kP_W1 = 0;
constexpr int kP_B1 = kP_W1 + kMlpH * kMlpIn;
constexpr int kP_W2 = kP_B1 + kMlpH;
constexpr int kP_B2 = kP_W2 + kMlpH * kMlpH;
constexpr int kP_W3 = kP_B2 + kMlpH;
constexpr int kP_B3 = kP_W3 + kMlpH;
constexpr int kNumParams = kP_B3 + 1;
// Gradient partials, summed in a fixed order by k_mlp_reduce (no float atomics: a run's training is the same
// every time).  Per wavefront of k_mlp_bwd_rows: b1, b2, w3, b3 and the loss; per workgroup of k_mlp_wgrad:
// W2 then W1.
constexpr int kRowB1 = 0, kRowB2 = kMlpH, kRowW3 = 2 * kMlpH, kRowB3 = 3 * kMlpH, kRowLoss = 3 * kMlpH + 1, kRowP = 3 * kMlpH + 2;
constexpr int kWgW2 = 0, kWgW1 = kMlpH * kMlpH, kWgP = kMlpH * kMlpH + kMlpH * kMlpIn;
constexpr int kWgRowsPerBlock = 256;

__device__ inline unsigned short f2bf(float f) {  // round to nearest even
    uint u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((uint)b << 16); }
using mlpd::silu;
using mlpd::mfma;
using mlpd::ld_bf4;
using mlpd::pack4_hw;
__device__ inline float dsilu(float z) {
    float s = __fdividef(1.0f, 1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

// Standardised input fragment of 16 rows starting at r0 (A layout).
__device__ inline bf16x4 load_x(const float* __restrict__ X, uint r0, uint n, const MlpDev& m) {
    const int l = threadIdx.x & 63, i = l & 15, kb = l >> 4;
    uint r = r0 + i;
    float v[4] = {0, 0, 0, 0};
    if (r < n) {
        float4 x = *(const float4*)(X + (size_t)r * kMlpIn + 4 * kb);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        int f = 4 * kb + e;
        v[e] = (r < n) ? (v[e] - m.in_mu[f]) * m.in_isd[f] : 0.0f;
    }
    return pack4(v[0], v[1], v[2], v[3]);
}

// Forward of 16 rows: returns the per-row output in lanes with l % 16 == 0 (rows 4 (l / 16) + e in y4).
// lds: 16 x 64 bf16 scratch of this wave.  z1/z2 (optional) receive the pre-activations.
__device__ inline f32x4 forward16(bf16x4 a1, const MlpDev& m, unsigned short* lds, float* z1, float* z2, uint r0, uint n) {
    const int l = threadIdx.x & 63, j = l & 15, kb = l >> 4, rb = l >> 4;
    // layer 1
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 d = {0, 0, 0, 0};
        d = mfma(a1, ld_bf4(m.w1 + (16 * nt + j) * kMlpIn + 4 * kb), d);
        float b = m.b1[16 * nt + j];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float z = d[e] + b;
            int row = 4 * rb + e;
            if (z1 && r0 + row < n) z1[(size_t)(r0 + row) * kMlpH + 16 * nt + j] = z;
            lds[row * kMlpH + 16 * nt + j] = f2bf(silu(z));
        }
    }
    __syncthreads();
    // layer 2 + output
    f32x4 p = {0, 0, 0, 0};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = mfma(ld_bf4(lds + (l & 15) * kMlpH + 16 * ks + 4 * kb), ld_bf4(m.w2 + (16 * nt + j) * kMlpH + 16 * ks + 4 * kb), acc);
        float b = m.b2[16 * nt + j], w = m.w3[16 * nt + j];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float z = acc[e] + b;
            int row = 4 * rb + e;
            if (z2 && r0 + row < n) z2[(size_t)(r0 + row) * kMlpH + 16 * nt + j] = z;
            p[e] += silu(z) * w;
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1)
#pragma unroll
        for (int e = 0; e < 4; ++e) p[e] += __shfl_xor(p[e], off, 64);
    float b3 = m.b3[0];
#pragma unroll
    for (int e = 0; e < 4; ++e) p[e] += b3;
    return p;
}


// Inference in transposed form: with the weights as the A operand and the activations as the B operand,
// H^T = W X^T leaves lane l holding H[row l % 16][units 4 (l / 16) .. + 3] of a 16-unit tile, which is the B
// fragment of the next layer's 16-deep K step.  Every lane keeps the weight fragments in registers; biases and
// the output weights sit in LDS (read as float4 per 4 units).  A wave walks 16-row tiles with a grid stride and
// loads the next tile's features before computing the current one.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ X, uint n, MlpDev m, float* __restrict__ y,
                                                   float4* __restrict__ xl, float w_scale, float kappa) {
    __shared__ __align__(16) float sb[3 * kMlpH];  // b1, b2, w3
    constexpr int kS = kMlpH + 8;  // padded row: the 16 rows a fragment read touches fall in different banks
    __shared__ __align__(16) unsigned short sw2[kMlpH * kS];  // W2, bf16 [unit][input]
    for (int t = threadIdx.x; t < 3 * kMlpH; t += blockDim.x)
        sb[t] = t < kMlpH ? m.b1[t] : (t < 2 * kMlpH ? m.b2[t - kMlpH] : m.w3[t - 2 * kMlpH]);
    for (int t = threadIdx.x; t < kMlpH * kMlpH / 4; t += blockDim.x) {
        const int row = t / (kMlpH / 4), c4 = t % (kMlpH / 4);
        ((uint2*)(sw2 + row * kS))[c4] = ((const uint2*)m.w2)[t];
    }
    __syncthreads();
    const int l = threadIdx.x & 63, i = l & 15, kb = l >> 4;
    bf16x4 w1[4];
    float mu[4], isd[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w1[t] = ld_bf4(m.w1 + (16 * t + i) * kMlpIn + 4 * kb);  // A = W1[16t + i][4kb + e]
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        mu[e] = m.in_mu[4 * kb + e];
        isd[e] = m.in_isd[4 * kb + e];
    }
    const float b3 = m.b3[0];
    const float4* sb4 = (const float4*)sb;  // sb4[4t + kb] = units 16t + 4kb .. + 3
    const uint waves = gridDim.x * (blockDim.x >> 6);
    uint tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    auto load = [&](uint tl) {
        const uint r = tl * 16 + i;
        return r < n ? *(const float4*)(X + (size_t)r * kMlpIn + 4 * kb) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    float4 x = load(tile);
    for (; (size_t)tile * 16 < n; tile += waves) {
        const uint r = tile * 16 + i;
        const float4 xn = load(tile + waves);  // the next tile's features, in flight during this tile
        // an offset the compiler cannot see through keeps the bias reads in the loop (hoisted, they would hold
        // 48 registers across it)
        uint o = (uint)kb;
        asm volatile("" : "+v"(o));
        const bf16x4 xb = pack4_hw((x.x - mu[0]) * isd[0], (x.y - mu[1]) * isd[1], (x.z - mu[2]) * isd[2],
                                   (x.w - mu[3]) * isd[3]);  // B = X^T[4kb + e][row i]; rows past n are 0
        bf16x4 h1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 d = {0, 0, 0, 0};
            d = mfma(w1[t], xb, d);  // D[unit 4kb + e][row i] of units 16t..16t+15
            const float4 b = sb4[4 * t + o];
            h1[t] = pack4_hw(silu(d[0] + b.x), silu(d[1] + b.y), silu(d[2] + b.z), silu(d[3] + b.w));
        }
        float p = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)  // A = W2[16t + i][16ks + 4kb + e] from LDS
                acc = mfma(*(const bf16x4*)(sw2 + (16 * t + i) * kS + 16 * ks + 4 * o), h1[ks], acc);
            const float4 b = sb4[16 + 4 * t + o], w = sb4[32 + 4 * t + o];
            p += silu(acc[0] + b.x) * w.x + silu(acc[1] + b.y) * w.y + silu(acc[2] + b.z) * w.z + silu(acc[3] + b.w) * w.w;
        }
        p += __shfl_xor(p, 16, 64);
        p += __shfl_xor(p, 32, 64);
        if (kb == 0 && r < n) {
            float out = (p + b3) * m.out_scale + m.out_bias;
            if (y) y[r] = out;
            if (xl) xl[r].w = warm_clamp(out * w_scale, xl[r].w, kappa);  // xl.w held lambda_0
        }
        x = xn;
    }
}

// Training forward over a batch: stores z1, z2 and y (standardised units).
__global__ __launch_bounds__(256) void k_mlp_fwd_train(const float* __restrict__ X, uint n, MlpDev m, float* __restrict__ z1,
                                                       float* __restrict__ z2, float* __restrict__ y) {
    __shared__ __align__(16) unsigned short lds[4][16 * kMlpH];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    uint r0 = (blockIdx.x * 4 + wave) * 16;
    bf16x4 a1 = load_x(X, r0, n, m);
    f32x4 p = forward16(a1, m, lds[wave], z1, z2, r0, n);
    if ((l & 15) == 0)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint r = r0 + 4 * (l >> 4) + e;
            if (r < n) y[r] = p[e];
        }
}

// Backward over rows: dy = (y - t) / n, dz2 = dy w3 silu'(z2), dh1 = dz2 W2 (MFMA), dz1 = dh1 silu'(z1).
// Writes this wavefront's db1, db2, dw3, db3 and loss to its partial (rowp + wavefront * kRowP); stores dz1, dz2
// for the weight gradients.
__global__ __launch_bounds__(256) void k_mlp_bwd_rows(const float* __restrict__ t, const float* __restrict__ y, const float* __restrict__ z1,
                                                      const float* __restrict__ z2, uint n, MlpDev m, float* __restrict__ dz1,
                                                      float* __restrict__ dz2, float* __restrict__ rowp) {
    __shared__ __align__(16) unsigned short lds[4][16 * kMlpH];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, j = l & 15, kb = l >> 4, rb = l >> 4;
    uint r0 = (blockIdx.x * 4 + wave) * 16;
    float* part = rowp + (size_t)(blockIdx.x * 4 + wave) * kRowP;
    const float inv_n = 1.0f / (float)n;
    float dy[4];
    float lsum = 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        uint r = r0 + 4 * rb + e;
        float d = (r < n) ? (y[r] - t[r]) : 0.0f;
        if (j == 0) lsum += 0.5f * d * d;  // each row once (16 lanes hold the same rows)
        dy[e] = d * inv_n;
    }
    float db3 = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        int o = 16 * nt + j;
        float w3 = m.w3[o];
        float gb2 = 0.0f, gw3 = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int row = 4 * rb + e;
            uint r = r0 + row;
            float dz = 0.0f;
            if (r < n) {
                float z = z2[(size_t)r * kMlpH + o];
                dz = dy[e] * w3 * dsilu(z);
                dz2[(size_t)r * kMlpH + o] = dz;
                gw3 += dy[e] * silu(z);
            }
            gb2 += dz;
            lds[wave][row * kMlpH + o] = f2bf(dz);
        }
        // reduce over the 4 row groups (lanes j, j+16, j+32, j+48)
        gb2 += __shfl_xor(gb2, 16, 64);
        gb2 += __shfl_xor(gb2, 32, 64);
        gw3 += __shfl_xor(gw3, 16, 64);
        gw3 += __shfl_xor(gw3, 32, 64);
        if (rb == 0) {
            part[kRowB2 + o] = gb2;
            part[kRowW3 + o] = gw3;
        }
    }
    if (j == 0)
#pragma unroll
        for (int e = 0; e < 4; ++e) db3 += dy[e];
    __syncthreads();
    // dh1[i][k] = sum_o dz2[i][o] W2[o][k]: A = dz2 (LDS), B[o][k] = W2T[k][o]
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = mfma(ld_bf4(lds[wave] + (l & 15) * kMlpH + 16 * ks + 4 * kb), ld_bf4(m.w2t + (16 * nt + j) * kMlpH + 16 * ks + 4 * kb), acc);
        int k = 16 * nt + j;
        float gb1 = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint r = r0 + 4 * rb + e;
            if (r < n) {
                float dz = acc[e] * dsilu(z1[(size_t)r * kMlpH + k]);
                dz1[(size_t)r * kMlpH + k] = dz;
                gb1 += dz;
            }
        }
        gb1 += __shfl_xor(gb1, 16, 64);
        gb1 += __shfl_xor(gb1, 32, 64);
        if (rb == 0) part[kRowB1 + k] = gb1;
    }
    // loss and db3: wave reduction
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lsum += __shfl_xor(lsum, off, 64);
        db3 += __shfl_xor(db3, off, 64);
    }
    if (l == 0) {
        part[kRowB3] = db3;
        part[kRowLoss] = lsum * inv_n;
    }
}

// Weight gradients dW2 = dz2^T h1 and dW1 = dz1^T x over a chunk of rows per block (MFMA, K = rows), staged
// through LDS as bf16.  Block: 4 waves; wave w owns output row tiles of dW2 (mt = w) and dW1 (mt = w).
constexpr int kWgRows = 64;
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ X, const float* __restrict__ z1,
                                                   const float* __restrict__ dz1, const float* __restrict__ dz2, uint n,
                                                   int rows_per_block, MlpDev m, float* __restrict__ wgp) {
    __shared__ __align__(16) unsigned short s_dz2[kWgRows * kMlpH], s_h1[kWgRows * kMlpH], s_dz1[kWgRows * kMlpH], s_x[kWgRows * kMlpIn];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, j = l & 15, kb = l >> 4;
    const int mt = wave;  // output rows 16 mt .. 16 mt + 15 (units o)
    f32x4 acc2[4], acc1 = {0, 0, 0, 0};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc2[nt] = f32x4{0, 0, 0, 0};
    uint rbeg = blockIdx.x * rows_per_block;
    uint rend = min(n, rbeg + rows_per_block);
    for (uint c0 = rbeg; c0 < rend; c0 += kWgRows) {
        __syncthreads();
        for (int t = threadIdx.x; t < kWgRows * kMlpH; t += 256) {
            int rr = t / kMlpH, cc = t % kMlpH;
            uint r = c0 + rr;
            bool ok = r < rend;
            s_dz2[t] = f2bf(ok ? dz2[(size_t)r * kMlpH + cc] : 0.0f);
            s_dz1[t] = f2bf(ok ? dz1[(size_t)r * kMlpH + cc] : 0.0f);
            s_h1[t] = f2bf(ok ? silu(z1[(size_t)r * kMlpH + cc]) : 0.0f);
        }
        for (int t = threadIdx.x; t < kWgRows * kMlpIn; t += 256) {
            int rr = t / kMlpIn, cc = t % kMlpIn;
            uint r = c0 + rr;
            s_x[t] = f2bf(r < rend ? (X[(size_t)r * kMlpIn + cc] - m.in_mu[cc]) * m.in_isd[cc] : 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kWgRows / 16; ++ks) {
            // A[o][r] = dz2[r][o] (o = 16 mt + l % 16, r = 16 ks + 4 kb + e)
            bf16x4 a2, a1;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int r = 16 * ks + 4 * kb + e;
                a2[e] = (short)s_dz2[r * kMlpH + 16 * mt + (l & 15)];
                a1[e] = (short)s_dz1[r * kMlpH + 16 * mt + (l & 15)];
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                bf16x4 b;
#pragma unroll
                for (int e = 0; e < 4; ++e) b[e] = (short)s_h1[(16 * ks + 4 * kb + e) * kMlpH + 16 * nt + j];
                acc2[nt] = mfma(a2, b, acc2[nt]);
            }
            bf16x4 bx;
#pragma unroll
            for (int e = 0; e < 4; ++e) bx[e] = (short)s_x[(16 * ks + 4 * kb + e) * kMlpIn + j];
            acc1 = mfma(a1, bx, acc1);
        }
    }
    // D[o][k]: lane holds rows o = 16 mt + 4 (l / 16) + e, column 16 nt + j
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        int o = 16 * mt + 4 * (l >> 4) + e;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) wgp[(size_t)blockIdx.x * kWgP + kWgW2 + o * kMlpH + 16 * nt + j] = acc2[nt][e];
        wgp[(size_t)blockIdx.x * kWgP + kWgW1 + o * kMlpIn + j] = acc1[e];
    }
}

// The gradient (and the loss at kNumParams) from the partials, each summed in partial order.
__global__ void k_mlp_reduce(const float* __restrict__ rowp, uint nrow, const float* __restrict__ wgp, uint nwg,
                             float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kNumParams) return;
    int q;
    bool rows = true;
    if (p < kP_B1) q = kWgW1 + (p - kP_W1), rows = false;
    else if (p < kP_W2) q = kRowB1 + (p - kP_B1);
    else if (p < kP_B2) q = kWgW2 + (p - kP_W2), rows = false;
    else if (p < kP_W3) q = kRowB2 + (p - kP_B2);
    else if (p < kP_B3) q = kRowW3 + (p - kP_W3);
    else if (p == kP_B3) q = kRowB3;
    else q = kRowLoss;
    float s = 0.0f;
    if (rows)
        for (uint b = 0; b < nrow; ++b) s += rowp[(size_t)b * kRowP + q];
    else
        for (uint b = 0; b < nwg; ++b) s += wgp[(size_t)b * kWgP + q];
    grad[p] = s;
}

// Adam on the FP32 master parameters, then the bf16 copies used by the MFMA kernels.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2, int np,
                           float lr, float b1, float b2, float eps, float c1, float c2, float wd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float gg = g[i] + wd * p[i];
    float a = b1 * m1[i] + (1.0f - b1) * gg;
    float b = b2 * m2[i] + (1.0f - b2) * gg * gg;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}
__global__ void k_mlp_pack(const float* __restrict__ p, unsigned short* __restrict__ w1, unsigned short* __restrict__ w2,
                           unsigned short* __restrict__ w2t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpH * kMlpIn) w1[i] = f2bf(p[kP_W1 + i]);
    if (i < kMlpH * kMlpH) {
        int o = i / kMlpH, k = i % kMlpH;
        unsigned short v = f2bf(p[kP_W2 + i]);
        w2[i] = v;
        w2t[k * kMlpH + o] = v;
    }
}

__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ T, const int* __restrict__ idx, int nb,
                              float* __restrict__ xb, float* __restrict__ tb) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nb) return;
    int s = idx[r];
    const float4* xs = (const float4*)(X + (size_t)s * kMlpIn);
    float4* xd = (float4*)(xb + (size_t)r * kMlpIn);
    for (int t = 0; t < 4; ++t) xd[t] = xs[t];
    tb[r] = T[s];
}

}  // namespace

// ============================================================================================ object
class rxMLP {
public:
    explicit rxMLP(unsigned seed) {
        m_ok = RX_HIP(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking)) &&
               RX_HIP(hipMalloc(&d_p, sizeof(float) * kNumParams)) && RX_HIP(hipMalloc(&d_g, sizeof(float) * (kNumParams + 1))) &&
               RX_HIP(hipMalloc(&d_m1, sizeof(float) * kNumParams)) && RX_HIP(hipMalloc(&d_m2, sizeof(float) * kNumParams)) &&
               RX_HIP(hipMalloc(&d_w1, sizeof(unsigned short) * kMlpH * kMlpIn)) &&
               RX_HIP(hipMalloc(&d_w2, sizeof(unsigned short) * kMlpH * kMlpH)) &&
               RX_HIP(hipMalloc(&d_w2t, sizeof(unsigned short) * kMlpH * kMlpH)) && RX_HIP(hipMalloc(&d_norm, sizeof(float) * 32));
        if (!m_ok) return;  // rxMLP_Create deletes the object
        // He-uniform initialisation (seeded); padded inputs 12..15 keep zero weights
        std::mt19937 g(seed);
        std::vector<float> p(kNumParams, 0.0f);
        auto uni = [&](float a) { return (float)((g() >> 8) * (1.0 / 16777216.0) * 2.0 - 1.0) * a; };
        for (int o = 0; o < kMlpH; ++o)
            for (int i = 0; i < 12; ++i) p[kP_W1 + o * kMlpIn + i] = uni(sqrtf(6.0f / 12.0f));
        for (int o = 0; o < kMlpH; ++o)
            for (int i = 0; i < kMlpH; ++i) p[kP_W2 + o * kMlpH + i] = uni(sqrtf(6.0f / 64.0f));
        for (int i = 0; i < kMlpH; ++i) p[kP_W3 + i] = uni(sqrtf(6.0f / 64.0f)) * 0.1f;
        set_params(p.data());
        float nrm[32];
        for (int f = 0; f < 16; ++f) {
            nrm[f] = 0.0f;
            nrm[16 + f] = f < 12 ? 1.0f : 0.0f;
        }
        set_norm(nrm, 1.0f, 0.0f);
    }
    ~rxMLP() {
        if (m_stream) RX_HIP(hipStreamSynchronize(m_stream));
        for (void* q : {(void*)d_p, (void*)d_g, (void*)d_m1, (void*)d_m2, (void*)d_w1, (void*)d_w2, (void*)d_w2t, (void*)d_norm})
            if (q) RX_HIP(hipFree(q));
        free_batch();
        if (m_stream) RX_HIP(hipStreamDestroy(m_stream));
    }
    void set_params(const float* h) {
        RX_HIP(hipMemcpyAsync(d_p, h, sizeof(float) * kNumParams, hipMemcpyHostToDevice, m_stream));
        rx_fill<float>(d_m1, 0.0f, kNumParams, m_stream);
        rx_fill<float>(d_m2, 0.0f, kNumParams, m_stream);
        m_t = 0;
        pack();
        RX_HIP(hipStreamSynchronize(m_stream));
    }
    void get_params(float* h) {
        RX_HIP(hipMemcpyAsync(h, d_p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, m_stream));
        RX_HIP(hipStreamSynchronize(m_stream));
    }
    void set_norm(const float* mu_isd, float out_scale, float out_bias) {
        RX_HIP(hipMemcpyAsync(d_norm, mu_isd, sizeof(float) * 32, hipMemcpyHostToDevice, m_stream));
        m_out_scale = out_scale;
        m_out_bias = out_bias;
        RX_HIP(hipStreamSynchronize(m_stream));
    }
    MlpDev view() const {
        MlpDev d;
        d.w1 = d_w1;
        d.w2 = d_w2;
        d.w2t = d_w2t;
        d.b1 = d_p + kP_B1;
        d.b2 = d_p + kP_B2;
        d.w3 = d_p + kP_W3;
        d.b3 = d_p + kP_B3;
        d.in_mu = d_norm;
        d.in_isd = d_norm + 16;
        d.out_scale = m_out_scale;
        d.out_bias = m_out_bias;
        return d;
    }
    void pack() { k_mlp_pack<<<rx_blocks(kMlpH * kMlpH, 256), 256, 0, m_stream>>>(d_p, d_w1, d_w2, d_w2t); }

    bool ensure_batch(uint b) {
        if (b <= m_bcap) return true;
        free_batch();
        const bool ok = RX_HIP(hipMalloc(&d_z1, sizeof(float) * kMlpH * b)) && RX_HIP(hipMalloc(&d_z2, sizeof(float) * kMlpH * b)) &&
                        RX_HIP(hipMalloc(&d_dz1, sizeof(float) * kMlpH * b)) && RX_HIP(hipMalloc(&d_dz2, sizeof(float) * kMlpH * b)) &&
                        RX_HIP(hipMalloc(&d_y, sizeof(float) * b)) && RX_HIP(hipMalloc(&d_xb, sizeof(float) * kMlpIn * b)) &&
                        RX_HIP(hipMalloc(&d_tb, sizeof(float) * b)) &&
                        RX_HIP(hipMalloc(&d_rowp, sizeof(float) * kRowP * 4 * ((b + 63) / 64))) &&
                        RX_HIP(hipMalloc(&d_wgp, sizeof(float) * kWgP * ((b + kWgRowsPerBlock - 1) / kWgRowsPerBlock)));
        if (!ok) {
            free_batch();
            return false;
        }
        m_bcap = b;
        return true;
    }
    void free_batch() {
        for (float** q : {&d_z1, &d_z2, &d_dz1, &d_dz2, &d_y, &d_xb, &d_tb, &d_rowp, &d_wgp})
            if (*q) {
                RX_HIP(hipFree(*q));
                *q = nullptr;
            }
        m_bcap = 0;
    }

    // Gradients of 0.5 mean (y - t)^2 over n rows (standardised target t) into d_g (+ loss at d_g[kNumParams]).
    bool gradients(const float* X, const float* T, uint n) {
        if (!ensure_batch(n)) return false;
        hipStream_t s = m_stream;
        MlpDev v = view();
        unsigned blocks = (n + 63) / 64;
        k_mlp_fwd_train<<<blocks, 256, 0, s>>>(X, n, v, d_z1, d_z2, d_y);
        k_mlp_bwd_rows<<<blocks, 256, 0, s>>>(T, d_y, d_z1, d_z2, n, v, d_dz1, d_dz2, d_rowp);
        const unsigned wblocks = (n + kWgRowsPerBlock - 1) / kWgRowsPerBlock;
        k_mlp_wgrad<<<wblocks, 256, 0, s>>>(X, d_z1, d_dz1, d_dz2, n, kWgRowsPerBlock, v, d_wgp);
        k_mlp_reduce<<<rx_blocks(kNumParams + 1, 256), 256, 0, s>>>(d_rowp, 4 * blocks, d_wgp, wblocks, d_g);
        return true;
    }

    void adam_step(float lr, float wd) {
        ++m_t;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = 1.0f / (1.0f - powf(b1, (float)m_t)), c2 = 1.0f / (1.0f - powf(b2, (float)m_t));
        k_mlp_adam<<<rx_blocks(kNumParams, 256), 256, 0, m_stream>>>(d_p, d_g, d_m1, d_m2, kNumParams, lr, b1, b2, 1e-8f, c1, c2, wd);
        pack();
    }

    int train(const float* hX, const float* hT, int n, int epochs, int batch, float lr, unsigned seed, float* hist) {
        hipStream_t s = m_stream;
        float *dX = nullptr, *dT = nullptr;
        int* didx = nullptr;
        auto release = [&]() {
            for (void* q : {(void*)dX, (void*)dT, (void*)didx})
                if (q) RX_HIP(hipFree(q));
        };
        if (!(RX_HIP(hipMalloc(&dX, sizeof(float) * kMlpIn * n)) && RX_HIP(hipMalloc(&dT, sizeof(float) * n)) &&
              RX_HIP(hipMalloc(&didx, sizeof(int) * n)) && ensure_batch((uint)batch))) {
            release();
            return 1;
        }
        RX_HIP(hipMemcpyAsync(dX, hX, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(dT, hT, sizeof(float) * n, hipMemcpyHostToDevice, s));
        std::mt19937 g(seed);
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        int nbatches = n / batch;
        for (int ep = 0; ep < epochs; ++ep) {
            for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[g() % (uint)(i + 1)]);
            RX_HIP(hipMemcpyAsync(didx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            float lr_e = lr * 0.5f * (1.0f + cosf(3.14159265f * ep / epochs));  // cosine schedule
            double eloss = 0.0;
            for (int b = 0; b < nbatches; ++b) {
                k_gather_rows<<<rx_blocks(batch, 256), 256, 0, s>>>(dX, dT, didx + b * batch, batch, d_xb, d_tb);
                gradients(d_xb, d_tb, batch);
                adam_step(lr_e, 0.0f);
                if (hist && (b == nbatches - 1)) {
                    float l = 0;
                    RX_HIP(hipMemcpyAsync(&l, d_g + kNumParams, sizeof(float), hipMemcpyDeviceToHost, s));
                    RX_HIP(hipStreamSynchronize(s));
                    eloss = l;
                }
            }
            if (hist) hist[ep] = (float)eloss;
        }
        RX_HIP(hipStreamSynchronize(s));
        release();
        return RX_HIP(hipGetLastError()) ? 0 : 1;
    }

    hipStream_t m_stream = nullptr;
    float *d_p = nullptr, *d_g = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_norm = nullptr;
    unsigned short *d_w1 = nullptr, *d_w2 = nullptr, *d_w2t = nullptr;
    float *d_z1 = nullptr, *d_z2 = nullptr, *d_dz1 = nullptr, *d_dz2 = nullptr, *d_y = nullptr, *d_xb = nullptr, *d_tb = nullptr;
    float *d_rowp = nullptr, *d_wgp = nullptr;  // gradient partials (per wavefront / per workgroup)
    uint m_bcap = 0;
    bool m_ok = false;  // every array of the constructor was allocated
    float m_out_scale = 1.0f, m_out_bias = 0.0f;
    int m_t = 0;
};

// 4 waves per block; up to 8 tiles of 16 rows per wave, at most 2048 blocks (8 per CU).
static unsigned infer_blocks(uint n) {
    uint tiles = (n + 15) / 16;
    uint blocks = (tiles + 31) / 32;
    return blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
}
void rx_mlp_infer_to_w(const float* feat, float4* xl, uint n, const MlpDev& m, float w_scale, float kappa, hipStream_t s) {
    if (n) k_mlp_infer<<<infer_blocks(n), 256, 0, s>>>(feat, n, m, nullptr, xl, w_scale, kappa);
}
void rx_mlp_infer(const float* feat, float* y, uint n, const MlpDev& m, hipStream_t s) {
    if (n) k_mlp_infer<<<infer_blocks(n), 256, 0, s>>>(feat, n, m, y, nullptr, 1.0f, 0.0f);
}
bool rx_mlp_view(void* obj, MlpDev* out) {
    if (!obj) return false;
    *out = ((rxMLP*)obj)->view();
    return true;
}

// Inference copy: bf16 W1 (2 KB) and W2 (8 KB), then FP32 b1, b2, w3, b3 and the standardisation constants.
namespace {
constexpr size_t kCopyW2 = sizeof(unsigned short) * kMlpH * kMlpIn;
constexpr size_t kCopyF = kCopyW2 + sizeof(unsigned short) * kMlpH * kMlpH;
constexpr size_t kCopyB1 = kCopyF, kCopyB2 = kCopyB1 + sizeof(float) * kMlpH;  // b2, w3, b3 follow contiguously
constexpr size_t kCopyNorm = kCopyB2 + sizeof(float) * (2 * kMlpH + 1);
constexpr size_t kCopyBytes = kCopyNorm + sizeof(float) * 32;
}  // namespace
size_t rx_mlp_copy_bytes() { return kCopyBytes; }

bool rx_mlp_copy(void* obj, void* dst, hipStream_t s, MlpDev* out) {
    if (!obj || !dst) return false;
    rxMLP* m = (rxMLP*)obj;
    // the packing kernel of the last SetParams runs on the network's own stream: finish it before copying
    if (!RX_HIP(hipStreamSynchronize(m->m_stream))) return false;
    const MlpDev v = m->view();
    unsigned char* d = (unsigned char*)dst;
    const bool ok =
        RX_HIP(hipMemcpyAsync(d, v.w1, kCopyW2, hipMemcpyDeviceToDevice, s)) &&
        RX_HIP(hipMemcpyAsync(d + kCopyW2, v.w2, kCopyF - kCopyW2, hipMemcpyDeviceToDevice, s)) &&
        RX_HIP(hipMemcpyAsync(d + kCopyB1, v.b1, sizeof(float) * kMlpH, hipMemcpyDeviceToDevice, s)) &&
        RX_HIP(hipMemcpyAsync(d + kCopyB2, v.b2, sizeof(float) * (2 * kMlpH + 1), hipMemcpyDeviceToDevice, s)) &&
        RX_HIP(hipMemcpyAsync(d + kCopyNorm, v.in_mu, sizeof(float) * 32, hipMemcpyDeviceToDevice, s));
    if (!ok) return false;
    MlpDev c;
    c.w1 = (const unsigned short*)d;
    c.w2 = (const unsigned short*)(d + kCopyW2);
    c.w2t = nullptr;  // backward only
    c.b1 = (const float*)(d + kCopyB1);
    c.b2 = (const float*)(d + kCopyB2);
    c.w3 = c.b2 + kMlpH;
    c.b3 = c.w3 + kMlpH;
    c.in_mu = (const float*)(d + kCopyNorm);
    c.in_isd = c.in_mu + 16;
    c.out_scale = v.out_scale;
    c.out_bias = v.out_bias;
    *out = c;
    return true;
}

extern "C" {

int rxMLP_NumParams(void) { return kNumParams; }
void* rxMLP_Create(unsigned seed) {
    rxMLP* m = new rxMLP(seed);
    if (!m->m_ok) {
        delete m;
        return nullptr;
    }
    return m;
}
void rxMLP_Destroy(void* h) { delete (rxMLP*)h; }
int rxMLP_SetParams(void* h, const float* params) {
    if (!h || !params) return 1;
    ((rxMLP*)h)->set_params(params);
    return 0;
}
int rxMLP_GetParams(void* h, float* params) {
    if (!h || !params) return 1;
    ((rxMLP*)h)->get_params(params);
    return 0;
}
int rxMLP_SetNormalization(void* h, const float* in_mu, const float* in_sd, float out_scale, float out_bias) {
    if (!h || !in_mu || !in_sd) return 1;
    float nrm[32];
    for (int f = 0; f < 16; ++f) {
        nrm[f] = f < 12 ? in_mu[f] : 0.0f;
        nrm[16 + f] = (f < 12 && in_sd[f] > 0.0f) ? 1.0f / in_sd[f] : 0.0f;
    }
    ((rxMLP*)h)->set_norm(nrm, out_scale, out_bias);
    return 0;
}
int rxMLP_Forward(void* h, const float* X, int n, float* Y) {
    if (!h || !X || !Y || n <= 0) return 1;
    rxMLP* p = (rxMLP*)h;
    float *dX = nullptr, *dY = nullptr;
    if (!(RX_HIP(hipMalloc(&dX, sizeof(float) * kMlpIn * n)) && RX_HIP(hipMalloc(&dY, sizeof(float) * n)))) return 1;
    RX_HIP(hipMemcpyAsync(dX, X, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, p->m_stream));
    rx_mlp_infer(dX, dY, (uint)n, p->view(), p->m_stream);
    RX_HIP(hipMemcpyAsync(Y, dY, sizeof(float) * n, hipMemcpyDeviceToHost, p->m_stream));
    RX_HIP(hipStreamSynchronize(p->m_stream));
    RX_HIP(hipFree(dX));
    RX_HIP(hipFree(dY));
    return RX_HIP(hipGetLastError()) ? 0 : 1;
}
int rxMLP_Gradients(void* h, const float* X, const float* T, int n, float* grads, float* loss) {
    if (!h || !X || !T || n <= 0) return 1;
    rxMLP* p = (rxMLP*)h;
    float *dX = nullptr, *dT = nullptr;
    bool ok = RX_HIP(hipMalloc(&dX, sizeof(float) * kMlpIn * n)) && RX_HIP(hipMalloc(&dT, sizeof(float) * n));
    if (ok) {
        RX_HIP(hipMemcpyAsync(dX, X, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, p->m_stream));
        RX_HIP(hipMemcpyAsync(dT, T, sizeof(float) * n, hipMemcpyHostToDevice, p->m_stream));
        ok = p->gradients(dX, dT, (uint)n);
    }
    if (!ok) {
        RX_HIP(hipStreamSynchronize(p->m_stream));
        for (void* q : {(void*)dX, (void*)dT})
            if (q) RX_HIP(hipFree(q));
        return 1;
    }
    if (grads) RX_HIP(hipMemcpyAsync(grads, p->d_g, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, p->m_stream));
    if (loss) RX_HIP(hipMemcpyAsync(loss, p->d_g + kNumParams, sizeof(float), hipMemcpyDeviceToHost, p->m_stream));
    RX_HIP(hipStreamSynchronize(p->m_stream));
    RX_HIP(hipFree(dX));
    RX_HIP(hipFree(dT));
    return RX_HIP(hipGetLastError()) ? 0 : 1;
}
int rxMLP_Train(void* h, const float* X, const float* T, int n, int epochs, int batch, float lr, unsigned seed, float* loss_history) {
    if (!h || !X || !T || n <= 0 || batch <= 0 || batch > n || epochs <= 0) {
        rx_set_error(1, "rxMLP_Train: bad arguments");
        return 1;
    }
    return ((rxMLP*)h)->train(X, T, n, epochs, batch, lr, seed, loss_history);
}

}  // extern "C"

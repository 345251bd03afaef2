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

// Inference over n feature rows (16 floats each): xl[i].w = (y_i * out_scale + out_bias) * w_scale.
void rx_mlp_infer_to_w(const float* feat, float4* xl, uint n, const MlpDev& m, float w_scale, hipStream_t s);
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

// heat.hip -- LaunchHeatTransferGPU (HeatTransfer.cu:19-36, :44-204): heat exchange of the ice/fluid particles.
//
// Per call, in the reference's order:
//   air     heats_i += ht (T_air - T_i) clamp((20 - s_i) / 20, 0, 1)  for surface particles (s_i > 0)
//   diffuse dT_i = cffCnTd * sum_{j in list(i), j != i} dt (T_j - T_i) / max(rho_j, 0.05) W_spline(1, r)
//   melt    heats[m] += 300 for every entry m != -1 of meltPrtIndx[1 .. 999]
//   phase   latent/sensible heat state machine (ice -2, melting -1, freezing 1, water 2; 250 = melt point)
// The first two are fused into one pass (their outputs are disjoint); the melt entries touch arbitrary
// particles and the phase pass reads every heat, so they stay separate launches.  All arithmetic is FP32 with
// no contraction, in the reference's operation order.
//
// Differences from the reference, all at its undefined edges: every kernel checks its index against the
// particle count (the reference launches a rounded-up grid without a bound); melt entries outside
// [0, prtNum) are ignored and repeated entries add 300 each (atomics in place of a racing +=).
#include "rx_common.h"
#include "../../include/rx_ice_funcs.h"

namespace {

constexpr int kNeighMax = 30;        // NEIGHT_MAX (HeatTransfer.cu:15)
constexpr float kTempMax = 500.0f;   // TEMP_MAX
constexpr float kTempMin = 0.0f;     // TEMP_MIN
constexpr float kPi = 3.14159265359f;
constexpr float kLatent = 165.0f;    // LATENT_HEAT
constexpr int kMeltSlots = 1000;     // size of the melt index array (HeatTransfar.cpp:101)

// KernelSpline (HeatTransfer.cu:62-75)
__device__ inline float spline(float d, float r, float w) {
#pragma clang fp contract(off)
    float q = d / r;
    if (q >= 0.0f && q < 1.0f) return w * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    if (q >= 1.0f && q < 2.0f) return w * 0.25f * (2.0f - q) * (2.0f - q) * (2.0f - q);
    return 0.0f;
}

// UpdateAirAndParticle (HeatTransfer.cu:79-96) + UpdateParticleAndParticle (:100-123).
__global__ void k_heat_exchange(float* __restrict__ heats, const float* __restrict__ temps, float* __restrict__ dtemps,
                                const int* __restrict__ surf, const int* __restrict__ neighs, const float* __restrict__ dens,
                                float air, float ht, float dt, float radius, float td, float wspline, int n) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ti = temps[i];
    int s = surf[i];
    if (s > 0) {
        float air_num = 20.0f - (float)s;
        if (air_num < 0) air_num = 0.0f;
        float area = air_num / 20.0f;
        if (area < 0.0f) area = 0.0f;
        if (area > 1.0f) area = 1.0f;
        heats[i] += ht * (air - ti) * area;
    }
    const int* nb = neighs + (size_t)i * kNeighMax;
    int cnt = nb[kNeighMax - 1];
    // the distance is fixed to 1 in the reference, so the kernel value is one constant per call
    const float kw = spline(1.0f, radius, wspline);
    float acc = 0.0f;
    for (int t = 0; t < cnt && t < kNeighMax; ++t) {
        int j = nb[t];
        if (j == i) continue;
        float d = dens[j];
        if (d < 0.05f) d = 0.05f;
        acc += dt * 1.0f * (temps[j] - ti) / d * kw;
    }
    dtemps[i] = acc * td;
}

// UpdateMeltParticle (HeatTransfer.cu:127-137)
__global__ void k_heat_melt(const int* __restrict__ melt, float* __restrict__ heats, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 || k >= kMeltSlots || k >= n) return;
    int m = melt[k];
    if (m < 0 || m >= n) return;
    atomicAdd(&heats[m], 300.0f);
}

// UpdatePhase (HeatTransfer.cu:140-204)
__global__ void k_heat_phase(float* __restrict__ heats, float* __restrict__ temps, const float* __restrict__ dtemps,
                             int* __restrict__ phase, int* __restrict__ changed, int n) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ph = phase[i];
    float t = temps[i], q = heats[i];
    if (ph == -2 && t > 250.0f) {
        ph = -1;
        t = 250.0f;
        q = 0.0f;
    } else if (ph == 2 && t < 250.0f) {
        ph = 1;
        t = 250.0f;
        q = kLatent;
    }
    if (ph == -1 || ph == 1) {
        float c = 2.1f + (2.1f * q / kLatent);
        q += dtemps[i] * c * 1.0f;
        if (ph == -1 && q < -kLatent / 300.0f) {
            ph = -2;
            t = 249.0f;
            q = 0.0f;
        } else if (ph == 1 && q > kLatent) {
            ph = 2;
            t = 251.0f;
            q = 0.0f;
        }
        if (ph == -1 && q > kLatent) {
            ph = 2;
            t = 251.0f;
            q = 0.0f;
            changed[i] = 1;
        } else if (ph == 1 && q < 0.0f) {
            ph = -2;
            t = 249.0f;
            q = 0.0f;
            changed[i] = 1;
        }
    } else {
        float c = (t > 250.0f) ? 4.2f : 2.1f;
        t = t + q / (c * 1.0f);
        q = 0.0f;
        t += dtemps[i];
        if (t > kTempMax) t = kTempMax;
        if (t < kTempMin) t = kTempMin;
    }
    phase[i] = ph;
    temps[i] = t;
    heats[i] = q;
}

}  // namespace

void LaunchHeatTransferGPU(float* heats, float* temps, float* dtemps, const int* surfParticles, const int* neightsParticles,
                           const float* densty, int* phase, int* phaseChangef, const int* meltPrtIndx, float radius, float airTemp,
                           float cffCnHt, float cffCnTd, float dt, int prtNum) {
    if (prtNum <= 0) return;
    if (!heats || !temps || !dtemps || !surfParticles || !neightsParticles || !densty || !phase || !phaseChangef) {
        rx_set_error(1, "LaunchHeatTransferGPU: null array");
        return;
    }
    hipStream_t s = rx_stream();
    // Wspline = 10 / (7 pi r^2) (HeatTransfer.cu:47), in the reference's float/double mix
    float wspline = 10.0f / (7.0f * kPi * (float)pow((double)radius, 2.0));
    unsigned nb = rx_blocks(prtNum, 256);
    k_heat_exchange<<<nb, 256, 0, s>>>(heats, temps, dtemps, surfParticles, neightsParticles, densty, airTemp, cffCnHt, dt, radius,
                                       cffCnTd, wspline, prtNum);
    if (meltPrtIndx) k_heat_melt<<<rx_blocks(kMeltSlots, 256), 256, 0, s>>>(meltPrtIndx, heats, prtNum);
    k_heat_phase<<<nb, 256, 0, s>>>(heats, temps, dtemps, phase, phaseChangef, prtNum);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));  // the reference synchronises before returning (HeatTransfer.cu:56)
}

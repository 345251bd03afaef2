// coupled.hip -- fluid + melting ice in one step loop (the ice mode frame of rxFlGLWindow::Idle,
// rx_fltk_glcanvas.cpp:1134-1145: fluid step, ice object step, heat transfer, ice structure).
//
// Every particle (fluid first, then the ice body) is a PBF particle of one rxPBDSPH_GPU with a thermal state.
// Per step:
//   1. the PBF step over all particles, followed by the heat exchange and phase update at the final positions;
//   2. gather of the ice particles (original index >= n_fluid) from the cell-ordered PBF state into the batched
//      shape-matching object, with the melt mask ice = phase < 0 (ice or melting);
//   3. the shape-matching substep (prediction from the post-fluid state, as OrientedPrtObj::UpdateCluster_SM
//      predicts from the SPH particles, ElasticObject_OP.cpp:572-608; clusters of ice members only);
//   4. write-back of the still-frozen particles into the PBF state (LaunchInterPolationGPU's blend of 1,
//      IceObject.cpp:876-897).  The write-back raises the PBF solver's largest-displacement bound, and the
//      next step starts from a grid built on the written-back positions.
// The two objects run on their own streams, ordered by events.
#include <vector>

#include "pbf_solver.h"
#include "../../include/rx_ice_funcs.h"
#include "../../include/rx_sph_host.h"

bool rx_ice_device_view(void* h, float4** x, float4** v, int** act, hipStream_t* s);

namespace {

__global__ void k_ice_gather(const float4* __restrict__ x, const float4* __restrict__ v, const uint* __restrict__ id,
                             const float4* __restrict__ therm, uint n, uint nf, float4* __restrict__ sx, float4* __restrict__ sv,
                             int* __restrict__ act) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint o = id[i];
    if (o < nf) return;
    uint k = o - nf;
    sx[k] = x[i];
    sv[k] = v[i];
    act[k] = therm[i].z < 0.0f ? 1 : 0;  // ice (-2) or melting (-1)
}

__global__ __launch_bounds__(256) void k_ice_scatter(float4* __restrict__ x, float4* __restrict__ v, const uint* __restrict__ id,
                                                     const float4* __restrict__ xb, uint n, uint nf, const float4* __restrict__ sx,
                                                     const float4* __restrict__ sv, const int* __restrict__ act,
                                                     float* __restrict__ disp2) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (i < n) {
        uint o = id[i];
        if (o >= nf && act[o - nf]) {
            float4 p = sx[o - nf];
            x[i] = make_float4(p.x, p.y, p.z, 0.0f);
            v[i] = sv[o - nf];
            float4 b = xb[i];
            float dx = p.x - b.x, dy = p.y - b.y, dz = p.z - b.z;
            m = dx * dx + dy * dy + dz * dz;
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax((unsigned int*)disp2, __float_as_uint(m));
}

}  // namespace

class rxCoupledGPU {
public:
    ~rxCoupledGPU() { release(); }

    int init(const rxSphEnv& fenv, const rxIceEnv& ienv, const float* pos4, const float* vel4, int nf, int ni, const float* temps,
             const int* phase) {
        if (nf < 0 || ni <= 0) {
            rx_set_error(1, "rxCoupled_Create: need n_fluid >= 0 and n_ice > 0");
            return 1;
        }
        m_nf = (uint)nf;
        m_ni = (uint)ni;
        m_pbf = new rxPBDSPH_GPU();
        if (m_pbf->Initialize(fenv) || m_pbf->SetParticles(pos4, vel4, nf + ni)) return 1;
        // the ice write-back moves particles by up to about h/2 per step: a wider skin and list keep the
        // corrections on the list path
        if (m_pbf->SetOption(RX_PBF_OPT_MAX_NEIGHBORS, 128.0f) || m_pbf->SetOption(RX_PBF_OPT_SKIN, 0.25f)) return 1;
        if (temps && m_pbf->SetThermal(temps, nullptr, phase)) return 1;
        m_thermal = temps != nullptr;
        m_sm = rxIceSM_Create(&ienv, pos4 + 4 * (size_t)nf, ni);  // rest shape: the initial ice positions
        if (!m_sm) return 1;
        if (!rx_ice_device_view(m_sm, &m_sx, &m_sv, &m_act, &m_sm_stream)) return 1;
        if (!(RX_HIP(hipEventCreateWithFlags(&m_ev_a, hipEventDisableTiming)) && RX_HIP(hipEventCreateWithFlags(&m_ev_b, hipEventDisableTiming))))
            return 1;
        return 0;
    }

    void release() {
        if (m_pbf) {
            m_pbf->Sync();
            delete m_pbf;
            m_pbf = nullptr;
        }
        if (m_sm) {
            rxIceSM_Destroy(m_sm);
            m_sm = nullptr;
        }
        if (m_ev_a) RX_HIP(hipEventDestroy(m_ev_a));
        if (m_ev_b) RX_HIP(hipEventDestroy(m_ev_b));
        m_ev_a = m_ev_b = nullptr;
    }

    int update(float dt, int steps) {
        if (!m_thermal) {
            rx_set_error(1, "rxCoupled_Update: the coupled step needs temperatures (melt mask)");
            return 1;
        }
        for (int t = 0; t < steps; ++t) {
            const int rc = m_pbf->Update(dt, 1);
            if (rc) return rc;
            rxPBDSPH_GPU::DeviceView d = m_pbf->device_view();
            uint n = m_pbf->size();
            unsigned nb = rx_blocks(n, 256);
            hipStream_t a = m_pbf->stream(), b = m_sm_stream;
            RX_HIP(hipEventRecord(m_ev_a, a));
            RX_HIP(hipStreamWaitEvent(b, m_ev_a, 0));
            k_ice_gather<<<nb, 256, 0, b>>>(d.x, d.v, d.id, d.therm, n, m_nf, m_sx, m_sv, m_act);
            if (rxIceSM_Update(m_sm, 1)) return 1;
            k_ice_scatter<<<nb, 256, 0, b>>>(d.x, d.v, d.id, d.xb, n, m_nf, m_sx, m_sv, m_act, d.stats + d.disp2_slot);
            RX_HIP(hipEventRecord(m_ev_b, b));
            RX_HIP(hipStreamWaitEvent(a, m_ev_b, 0));
            m_pbf->invalidate_list();  // the next prediction pass walks a grid built on the written-back state
            if (!RX_HIP(hipGetLastError())) return 1;
        }
        return 0;
    }

    rxPBDSPH_GPU* m_pbf = nullptr;
    void* m_sm = nullptr;

private:
    uint m_nf = 0, m_ni = 0;
    bool m_thermal = false;
    float4 *m_sx = nullptr, *m_sv = nullptr;
    int* m_act = nullptr;
    hipStream_t m_sm_stream = nullptr;
    hipEvent_t m_ev_a = nullptr, m_ev_b = nullptr;
};

extern "C" {

void* rxCoupled_Create(const rxSphEnv* fluid_env, const rxIceEnv* ice_env, const float* pos4, const float* vel4, int n_fluid, int n_ice,
                       const float* temps, const int* phase) {
    if (!fluid_env || !ice_env || !pos4) {
        rx_set_error(1, "rxCoupled_Create: null argument");
        return nullptr;
    }
    rxCoupledGPU* c = new rxCoupledGPU();
    if (c->init(*fluid_env, *ice_env, pos4, vel4, n_fluid, n_ice, temps, phase)) {
        delete c;
        return nullptr;
    }
    return c;
}
void rxCoupled_Destroy(void* h) { delete (rxCoupledGPU*)h; }
int rxCoupled_Update(void* h, float dt, int steps) { return h ? ((rxCoupledGPU*)h)->update(dt, steps) : 1; }
int rxCoupled_Sync(void* h) {
    if (!h) return 1;
    rxCoupledGPU* c = (rxCoupledGPU*)h;
    return c->m_pbf->Sync();
}
int rxCoupled_GetArray(void* h, int type, float* host, int count) { return h ? ((rxCoupledGPU*)h)->m_pbf->GetArray(type, host, count) : 1; }
int rxCoupled_GetStats(void* h, float* out, int count) { return h ? ((rxCoupledGPU*)h)->m_pbf->GetStats(out, count) : 1; }
int rxCoupled_SetOption(void* h, int option, float value) { return h ? ((rxCoupledGPU*)h)->m_pbf->SetOption(option, value) : 1; }
int rxCoupled_GetProfile(void* h, float* ms, int* count, int n) { return h ? ((rxCoupledGPU*)h)->m_pbf->GetProfile(ms, count, n) : 1; }
int rxCoupled_SetNetwork(void* h, void* mlp) { return h ? ((rxCoupledGPU*)h)->m_pbf->SetNetwork(mlp) : 1; }
int rxCoupled_SetIceTiles(void* h, int autotune, int* out_tiles) {
    if (!h) return 1;
    rxCoupledGPU* c = (rxCoupledGPU*)h;
    c->m_pbf->Sync();
    return autotune ? rxIceSM_Autotune(c->m_sm, out_tiles) : 0;
}

}  // extern "C"

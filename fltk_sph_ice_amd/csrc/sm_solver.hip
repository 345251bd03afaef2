// sm_solver.hip -- batched ice shape matching (semantics in include/rx_ice_funcs.h).
//
// Cluster pass: one persistent launch over work units; a unit is 64/T clusters of one size bin, each handled
// by a sub-group of T lanes (T = 8, 16, 32 or 64, chosen per bin by the autotuner).  A sub-group accumulates
// its members relative to the cluster's own particle (sum d, sum q, sum d q^T), reduces with xor shuffles,
// and every lane of the sub-group computes the polar decomposition (Jacobi on A^T A, ShapeMatching.h:242-387)
// redundantly.  The cluster writes one affine goal transform (RL, t): goal(p) = RL rest_p + t.
// Averaging pass: each particle gathers the transforms of the clusters it belongs to (48 B each, L2
// resident) -- no per-member positions are ever written.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include "rx_common.h"
#include "../../include/rx_ice_funcs.h"

namespace {

constexpr int kNumBins = 5;
constexpr int kBinMax[kNumBins] = {8, 16, 32, 64, 1 << 30};

struct SmK {
    float alpha, beta, dt, res;
    float3 g;
    float3 lo, hi;
    int volcon;
    uint n;
    int masked;  // a melt mask is installed (active[] decides which particles and clusters take part)
};

// ------------------------------------------------------------------------------------------ 3x3 helpers
// EigenJacobiMethod (ShapeMatching.h:242-343) for n = 3 in FP32: largest off-diagonal pivot, same rotation;
// stops when the off-diagonal magnitude is below 1e-7 of the trace (FP32 cannot reach the reference's
// absolute 1e-8 for all scales) or after 24 rotations.
__device__ inline void jacobi3(float a[9], float v[9]) {
    for (int it = 0; it < 24; ++it) {
        float a01 = fabsf(a[1]), a02 = fabsf(a[2]), a12 = fabsf(a[5]);
        int i = 0, j = 1;
        float x = a01;
        if (a02 > x) { i = 0; j = 2; x = a02; }
        if (a12 > x) { i = 1; j = 2; x = a12; }
        float tr = fabsf(a[0]) + fabsf(a[4]) + fabsf(a[8]);
        if (x == 0.0f || 2.0f * (a01 + a02 + a12) < 1e-7f * tr) break;
        float aii = a[i * 3 + i], ajj = a[j * 3 + j], aij = a[i * 3 + j];
        float al = 0.5f * (aii - ajj);
        float be = sqrtf(al * al + aij * aij);
        float ct = sqrtf(0.5f * (1.0f + fabsf(al) / be));
        float st = (((aii - ajj) >= 0.0f) ? 1.0f : -1.0f) * aij / (2.0f * be * ct);
        int m = 3 - i - j;  // the remaining index
        float aim = a[i * 3 + m], ajm = a[j * 3 + m];
        float bim = aim * ct + ajm * st;
        float bjm = -aim * st + ajm * ct;
        a[i * 3 + m] = a[m * 3 + i] = bim;
        a[j * 3 + m] = a[m * 3 + j] = bjm;
        a[i * 3 + i] = aii * ct * ct + 2.0f * aij * ct * st + ajj * st * st;
        a[j * 3 + j] = aii * st * st - 2.0f * aij * ct * st + ajj * ct * ct;
        a[i * 3 + j] = a[j * 3 + i] = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float vi = v[r * 3 + i], vj = v[r * 3 + j];
            v[r * 3 + i] = vi * ct + vj * st;
            v[r * 3 + j] = -vi * st + vj * ct;
        }
    }
}

// PolarDecomposition (ShapeMatching.h:352-387): R = A (A^T A)^(-1/2).
__device__ inline void polar3(const float A[9], float R[9]) {
    float ata[9], U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) ata[r * 3 + c] = A[0 * 3 + r] * A[0 * 3 + c] + A[1 * 3 + r] * A[1 * 3 + c] + A[2 * 3 + r] * A[2 * 3 + c];
    jacobi3(ata, U);
    float l[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = (ata[k * 4] <= 0.0f) ? 0.0f : 1.0f / sqrtf(ata[k * 4]);
    float S1[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) S1[r * 3 + c] = l[0] * U[r * 3] * U[c * 3] + l[1] * U[r * 3 + 1] * U[c * 3 + 1] + l[2] * U[r * 3 + 2] * U[c * 3 + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[r * 3 + c] = A[r * 3] * S1[c] + A[r * 3 + 1] * S1[3 + c] + A[r * 3 + 2] * S1[6 + c];
}

__device__ inline float det3(const float a[9]) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// ------------------------------------------------------------------------------------------ kernels
// Prediction + wall response of OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:572-608).
__global__ void k_sm_predict(const float4* __restrict__ x, const float4* __restrict__ v, float4* __restrict__ xs, SmK k) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.n) return;
    float p[3] = {x[i].x, x[i].y, x[i].z}, vv[3] = {v[i].x, v[i].y, v[i].z};
    float g[3] = {k.g.x, k.g.y, k.g.z}, lo[3] = {k.lo.x, k.lo.y, k.lo.z}, hi[3] = {k.hi.x, k.hi.y, k.hi.z};
    float c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = p[a] + (vv[a] + g[a] * k.dt) * k.dt;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] < lo[a] || c[a] > hi[a]) {
            c[a] = p[a] - (vv[a] + g[a] * k.dt) * k.dt * k.res;
#pragma unroll
            for (int o = 0; o < 3; ++o)
                if (o != a) c[o] = p[o];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = fminf(fmaxf(c[a], lo[a]), hi[a]);
    xs[i] = make_float4(c[0], c[1], c[2], 0.0f);
}

// 3x3 inverse (cofactors / determinant); false when the determinant is zero.
__device__ inline bool inv3f(const float a[9], float o[9]) {
    float d = det3(a);
    if (d == 0.0f) return false;
    float id = 1.0f / d;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

// One cluster per sub-group of T lanes.  Masked: only active members count; when some members are inactive the
// rest centre of mass and Aqq^-1 are those of the active members (from the same sums), and below 4 active
// members the linear term is dropped.
template <int T, bool Masked>
__device__ inline void sm_cluster(int c, bool valid, const int* __restrict__ cs, const int* __restrict__ mem, const float4* __restrict__ xs,
                                  const float4* __restrict__ rest, const float4* __restrict__ pre, float4* __restrict__ tr,
                                  const int* __restrict__ act, const SmK& k) {
    const int sub = threadIdx.x & (T - 1);
    constexpr int NA = Masked ? 22 : 15;
    int s = 0, e = 0;
    float3 xref = f3(0, 0, 0), cm0 = f3(0, 0, 0);
    if (valid && Masked && !act[c]) valid = false;
    if (valid) {
        s = cs[c];
        e = cs[c + 1];
        xref = f3(xs[c]);
        cm0 = f3(pre[3 * c]);
    }
    float acc[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) acc[t] = 0.0f;
    auto member = [&](bool on, const float4& xm, const float4& rm) {
        if (Masked && !on) return;
        float3 d = f3(xm) - xref;
        float3 q = f3(rm) - cm0;
        acc[0] += d.x; acc[1] += d.y; acc[2] += d.z;
        acc[3] += q.x; acc[4] += q.y; acc[5] += q.z;
        acc[6] += d.x * q.x; acc[7] += d.x * q.y; acc[8] += d.x * q.z;
        acc[9] += d.y * q.x; acc[10] += d.y * q.y; acc[11] += d.y * q.z;
        acc[12] += d.z * q.x; acc[13] += d.z * q.y; acc[14] += d.z * q.z;
        if (Masked) {
            acc[15] += 1.0f;
            acc[16 % NA] += q.x * q.x; acc[17 % NA] += q.x * q.y; acc[18 % NA] += q.x * q.z;
            acc[19 % NA] += q.y * q.y; acc[20 % NA] += q.y * q.z; acc[21 % NA] += q.z * q.z;
        }
    };
    int t = s + sub;
    for (; t + T < e; t += 2 * T) {  // two members in flight per lane
        const int m0 = mem[t], m1 = mem[t + T];
        const float4 x0 = xs[m0], r0 = rest[m0], x1 = xs[m1], r1 = rest[m1];
        const bool a0 = Masked ? act[m0] != 0 : true, a1 = Masked ? act[m1] != 0 : true;
        member(a0, x0, r0);
        member(a1, x1, r1);
    }
    if (t < e) {
        const int m = mem[t];
        member(Masked ? act[m] != 0 : true, xs[m], rest[m]);
    }
#pragma unroll
    for (int off = T / 2; off > 0; off >>= 1)
#pragma unroll
        for (int t = 0; t < NA; ++t) acc[t] += __shfl_xor(acc[t], off, 64);
    if (!valid) return;
    const int kall = e - s;
    const int kk = Masked ? (int)acc[15 % NA] : kall;
    float RL[9];
    float3 cm;
    if (kk <= 1) {  // rxShapeMatching::shapeMatching returns early (ShapeMatching.cpp:494): goal = x*
#pragma unroll
        for (int t = 0; t < 9; ++t) RL[t] = (t % 4 == 0) ? 1.0f : 0.0f;
        cm = xref;
        cm0 = f3(rest[c]);
    } else {
        float ik = 1.0f / (float)kk;
        float3 md = f3(acc[0] * ik, acc[1] * ik, acc[2] * ik);
        cm = xref + md;
        // Apq = sum (x - cm) q^T = sum d q^T - md (sum q)^T  (independent of the rest origin of q)
        float apq[9];
        float sq[3] = {acc[3], acc[4], acc[5]}, mdv[3] = {md.x, md.y, md.z};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) apq[r * 3 + cc] = acc[6 + r * 3 + cc] - mdv[r] * sq[cc];
        float R[9];
        polar3(apq, R);
        bool linear = k.beta != 0.0f;
        float ai[9];
        if (linear) {
            float4 p1 = pre[3 * c + 1], p2 = pre[3 * c + 2];
            ai[0] = pre[3 * c].w; ai[1] = p1.x; ai[2] = p1.y; ai[3] = p1.z; ai[4] = p1.w;
            ai[5] = p2.x; ai[6] = p2.y; ai[7] = p2.z; ai[8] = p2.w;
        }
        if (Masked && kk < kall) {
            // rest centre of mass of the active members and their Aqq = sum q q^T - k delta delta^T
            float3 dl = f3(sq[0] * ik, sq[1] * ik, sq[2] * ik);
            cm0 = cm0 + dl;
            if (kk < 4) {
                linear = false;
            } else if (linear) {
                const float kf = (float)kk;
                float aq[9] = {acc[16 % NA] - kf * dl.x * dl.x, acc[17 % NA] - kf * dl.x * dl.y, acc[18 % NA] - kf * dl.x * dl.z,
                               0.0f, acc[19 % NA] - kf * dl.y * dl.y, acc[20 % NA] - kf * dl.y * dl.z,
                               0.0f, 0.0f, acc[21 % NA] - kf * dl.z * dl.z};
                aq[3] = aq[1];
                aq[6] = aq[2];
                aq[7] = aq[5];
                if (!inv3f(aq, ai))
#pragma unroll
                    for (int t = 0; t < 9; ++t) ai[t] = 0.0f;
            }
        }
        if (linear) {
            float A[9];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) A[r * 3 + cc] = apq[r * 3] * ai[cc] + apq[r * 3 + 1] * ai[3 + cc] + apq[r * 3 + 2] * ai[6 + cc];
            if (k.volcon) {
                float d = fabsf(det3(A));
                if (d > 1e-10f) {
                    float sc = fminf(1.0f / sqrtf(d), 2.0f);
#pragma unroll
                    for (int t = 0; t < 9; ++t) A[t] *= sc;
                }
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) RL[t] = k.beta * A[t] + (1.0f - k.beta) * R[t];
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) RL[t] = R[t];
        }
    }
    if (sub == 0) {
        // goal(p) = RL (rest_p - cm0) + cm = RL rest_p + (cm - RL cm0)
        float3 t3 = cm - f3(RL[0] * cm0.x + RL[1] * cm0.y + RL[2] * cm0.z, RL[3] * cm0.x + RL[4] * cm0.y + RL[5] * cm0.z,
                            RL[6] * cm0.x + RL[7] * cm0.y + RL[8] * cm0.z);
        tr[3 * c + 0] = make_float4(RL[0], RL[1], RL[2], t3.x);
        tr[3 * c + 1] = make_float4(RL[3], RL[4], RL[5], t3.y);
        tr[3 * c + 2] = make_float4(RL[6], RL[7], RL[8], t3.z);
    }
}

template <int T>
__device__ inline void sm_cluster_sel(int c, bool valid, const int* __restrict__ cs, const int* __restrict__ mem, const float4* __restrict__ xs,
                                      const float4* __restrict__ rest, const float4* __restrict__ pre, float4* __restrict__ tr,
                                      const int* __restrict__ act, const SmK& k) {
    if (k.masked) sm_cluster<T, true>(c, valid, cs, mem, xs, rest, pre, tr, act, k);
    else sm_cluster<T, false>(c, valid, cs, mem, xs, rest, pre, tr, act, k);
}

// Persistent cluster pass: waves loop over work units (bin, first index into the bin-ordered cluster list).
__global__ __launch_bounds__(256) void k_sm_clusters(const int2* __restrict__ units, int nunits, const int* __restrict__ tiles,
                                                     const int* __restrict__ bin_end, const int* __restrict__ corder,
                                                     const int* __restrict__ cs, const int* __restrict__ mem,
                                                     const float4* __restrict__ xs, const float4* __restrict__ rest,
                                                     const float4* __restrict__ pre, float4* __restrict__ tr,
                                                     const int* __restrict__ act, SmK k) {
    const int wpb = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int u = (int)xcd_block(blockIdx.x, gridDim.x) * wpb + wave; u < nunits; u += gridDim.x * wpb) {
        int2 un = units[u];
        int T = tiles[un.x];
        int idx = un.y + lane / T;
        bool valid = idx < bin_end[un.x];
        int c = valid ? corder[idx] : 0;
        switch (T) {
        case 8: sm_cluster_sel<8>(c, valid, cs, mem, xs, rest, pre, tr, act, k); break;
        case 16: sm_cluster_sel<16>(c, valid, cs, mem, xs, rest, pre, tr, act, k); break;
        case 32: sm_cluster_sel<32>(c, valid, cs, mem, xs, rest, pre, tr, act, k); break;
        default: sm_cluster_sel<64>(c, valid, cs, mem, xs, rest, pre, tr, act, k); break;
        }
    }
}

// ConvoNormal::StepConvolution (Ice_Convolution_Normal.cpp:32-67) over the clusters containing each particle:
// x_p = x*_p + alpha (mean_c goal_c(p) - x*_p).
// With a melt mask, inactive particles keep x* and only active clusters count.
__global__ void k_sm_average(const int* __restrict__ ps, const int* __restrict__ pc, const float4* __restrict__ tr,
                             const float4* __restrict__ rest, const float4* __restrict__ xs, float4* __restrict__ xo,
                             const int* __restrict__ act, SmK k) {
    uint i = xcd_thread();
    if (i >= k.n) return;
    int s = ps[i], e = ps[i + 1];
    float3 x = f3(xs[i]);
    if (k.masked && !act[i]) {
        xo[i] = make_float4(x.x, x.y, x.z, 0.0f);
        return;
    }
    float3 r = f3(rest[i]);
    float3 acc = f3(0, 0, 0);
    int cnt = 0;
    auto add = [&](int c, const float4& a, const float4& b, const float4& d) {
        if (k.masked && !act[c]) return;
        acc += f3(a.x * r.x + a.y * r.y + a.z * r.z + a.w, b.x * r.x + b.y * r.y + b.z * r.z + b.w, d.x * r.x + d.y * r.y + d.z * r.z + d.w);
        ++cnt;
    };
    int t = s;
    for (; t + 3 < e; t += 4) {  // four clusters' transforms in flight before any is used
        const int c0 = pc[t], c1 = pc[t + 1], c2 = pc[t + 2], c3 = pc[t + 3];
        const float4 a0 = tr[3 * c0], b0 = tr[3 * c0 + 1], d0 = tr[3 * c0 + 2];
        const float4 a1 = tr[3 * c1], b1 = tr[3 * c1 + 1], d1 = tr[3 * c1 + 2];
        const float4 a2 = tr[3 * c2], b2 = tr[3 * c2 + 1], d2 = tr[3 * c2 + 2];
        const float4 a3 = tr[3 * c3], b3 = tr[3 * c3 + 1], d3 = tr[3 * c3 + 2];
        add(c0, a0, b0, d0);
        add(c1, a1, b1, d1);
        add(c2, a2, b2, d2);
        add(c3, a3, b3, d3);
    }
    for (; t < e; ++t) {
        const int c = pc[t];
        add(c, tr[3 * c], tr[3 * c + 1], tr[3 * c + 2]);
    }
    if (cnt > 0) x = x + k.alpha * ((1.0f / (float)cnt) * acc - x);
    xo[i] = make_float4(x.x, x.y, x.z, 0.0f);
}

// OrientedParticle::UpdatePosAndVel (OrientedParticle.cpp:377-393)
__global__ void k_sm_velocity(float4* __restrict__ x, float4* __restrict__ v, const float4* __restrict__ xn, SmK k) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.n) return;
    float3 a = f3(xn[i]), b = f3(x[i]);
    float3 vv = (1.0f / k.dt) * (a - b);
    v[i] = make_float4(vv.x, vv.y, vv.z, 0.0f);
    x[i] = make_float4(a.x, a.y, a.z, 0.0f);
}

int bin_of(int size) {
    for (int b = 0; b < kNumBins; ++b)
        if (size <= kBinMax[b]) return b;
    return kNumBins - 1;
}

}  // namespace

// ============================================================================================ object
class rxIceSM {
public:
    ~rxIceSM() { release(); }

    int init(const rxIceEnv& env, const float* rest4, int n) {
        m_env = env;
        m_n = (uint)n;
        if (!RX_HIP(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking))) return 1;
        build_clusters(rest4);
        return upload(rest4);
    }

    void release() {
        if (m_stream) RX_HIP(hipStreamSynchronize(m_stream));
        drop_graph();
        for (void** p : {(void**)&d_x, (void**)&d_v, (void**)&d_xa, (void**)&d_xb, (void**)&d_rest, (void**)&d_pre, (void**)&d_tr,
                         (void**)&d_cs, (void**)&d_mem, (void**)&d_ps, (void**)&d_pc, (void**)&d_units, (void**)&d_tiles,
                         (void**)&d_binend, (void**)&d_corder, (void**)&d_act})
            if (*p) {
                RX_HIP(hipFree(*p));
                *p = nullptr;
            }
        for (hipEvent_t e : m_ev) RX_HIP(hipEventDestroy(e));
        m_ev.clear();
        if (m_stream) {
            RX_HIP(hipStreamDestroy(m_stream));
            m_stream = nullptr;
        }
    }

    // Clusters: {p} and every particle within the radius of p (host hashed grid, members ascending).
    void build_clusters(const float* rest4) {
        const float R = m_env.cluster_radius;
        const double R2 = (double)R * R;
        float lo[3] = {1e30f, 1e30f, 1e30f};
        for (uint i = 0; i < m_n; ++i)
            for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], rest4[4 * i + a]);
        std::unordered_map<long long, std::vector<int>> cells;
        auto key = [&](int x, int y, int z) { return ((long long)x * 73856093LL) ^ ((long long)y * 19349663LL) ^ ((long long)z * 83492791LL); };
        std::vector<int> cx(m_n), cy(m_n), cz(m_n);
        for (uint i = 0; i < m_n; ++i) {
            cx[i] = (int)floor((rest4[4 * i] - lo[0]) / R);
            cy[i] = (int)floor((rest4[4 * i + 1] - lo[1]) / R);
            cz[i] = (int)floor((rest4[4 * i + 2] - lo[2]) / R);
            cells[key(cx[i], cy[i], cz[i])].push_back((int)i);
        }
        h_cs.assign(m_n + 1, 0);
        h_mem.clear();
        h_mem.reserve((size_t)m_n * 64);
        std::vector<int> tmp;
        for (uint i = 0; i < m_n; ++i) {
            tmp.clear();
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        auto it = cells.find(key(cx[i] + dx, cy[i] + dy, cz[i] + dz));
                        if (it == cells.end()) continue;
                        for (int j : it->second) {
                            if (cx[j] != cx[i] + dx || cy[j] != cy[i] + dy || cz[j] != cz[i] + dz) continue;  // hash collision
                            double ddx = (double)rest4[4 * i] - rest4[4 * j], ddy = (double)rest4[4 * i + 1] - rest4[4 * j + 1],
                                   ddz = (double)rest4[4 * i + 2] - rest4[4 * j + 2];
                            if (ddx * ddx + ddy * ddy + ddz * ddz <= R2) tmp.push_back(j);
                        }
                    }
            std::sort(tmp.begin(), tmp.end());
            h_mem.insert(h_mem.end(), tmp.begin(), tmp.end());
            h_cs[i + 1] = (int)h_mem.size();
        }
        // transpose: clusters containing each particle
        h_ps.assign(m_n + 1, 0);
        for (int m : h_mem) h_ps[m + 1]++;
        for (uint i = 0; i < m_n; ++i) h_ps[i + 1] += h_ps[i];
        h_pc.assign(h_mem.size(), 0);
        std::vector<int> fill(h_ps.begin(), h_ps.end() - 1);
        for (uint c = 0; c < m_n; ++c)
            for (int t = h_cs[c]; t < h_cs[c + 1]; ++t) h_pc[fill[h_mem[t]]++] = (int)c;
        // rest centre of mass and Aqq^-1 per cluster (double, unit masses)
        h_pre.assign((size_t)m_n * 12, 0.0f);
        for (uint c = 0; c < m_n; ++c) {
            int s = h_cs[c], e = h_cs[c + 1], kk = e - s;
            double cm[3] = {0, 0, 0};
            for (int t = s; t < e; ++t)
                for (int a = 0; a < 3; ++a) cm[a] += rest4[4 * h_mem[t] + a];
            for (int a = 0; a < 3; ++a) cm[a] /= std::max(kk, 1);
            double q[9] = {0};
            for (int t = s; t < e; ++t) {
                double d[3];
                for (int a = 0; a < 3; ++a) d[a] = rest4[4 * h_mem[t] + a] - cm[a];
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) q[r * 3 + cc] += d[r] * d[cc];
            }
            double inv[9] = {0};
            double det = q[0] * (q[4] * q[8] - q[5] * q[7]) - q[1] * (q[3] * q[8] - q[5] * q[6]) + q[2] * (q[3] * q[7] - q[4] * q[6]);
            if (det != 0.0) {
                double id = 1.0 / det;
                inv[0] = (q[4] * q[8] - q[5] * q[7]) * id;
                inv[1] = (q[2] * q[7] - q[1] * q[8]) * id;
                inv[2] = (q[1] * q[5] - q[2] * q[4]) * id;
                inv[3] = (q[5] * q[6] - q[3] * q[8]) * id;
                inv[4] = (q[0] * q[8] - q[2] * q[6]) * id;
                inv[5] = (q[2] * q[3] - q[0] * q[5]) * id;
                inv[6] = (q[3] * q[7] - q[4] * q[6]) * id;
                inv[7] = (q[1] * q[6] - q[0] * q[7]) * id;
                inv[8] = (q[0] * q[4] - q[1] * q[3]) * id;
            }
            float* o = &h_pre[(size_t)c * 12];
            o[0] = (float)cm[0];
            o[1] = (float)cm[1];
            o[2] = (float)cm[2];
            for (int t = 0; t < 9; ++t) o[3 + t] = (float)inv[t];
        }
        // bins
        h_corder.resize(m_n);
        std::vector<int> cnt(kNumBins, 0);
        for (uint c = 0; c < m_n; ++c) cnt[bin_of(h_cs[c + 1] - h_cs[c])]++;
        h_binstart.assign(kNumBins + 1, 0);
        for (int b = 0; b < kNumBins; ++b) h_binstart[b + 1] = h_binstart[b] + cnt[b];
        std::vector<int> pos(h_binstart.begin(), h_binstart.end() - 1);
        for (uint c = 0; c < m_n; ++c) h_corder[pos[bin_of(h_cs[c + 1] - h_cs[c])]++] = (int)c;
        for (int b = 0; b < kNumBins; ++b) m_tiles[b] = std::min(64, std::max(8, kBinMax[b] >= (1 << 29) ? 64 : kBinMax[b]));
    }

    void make_units(std::vector<int2>& units, int only_bin = -1) const {
        units.clear();
        for (int b = 0; b < kNumBins; ++b) {
            if (only_bin >= 0 && b != only_bin) continue;
            int per = 64 / m_tiles[b];
            for (int i = h_binstart[b]; i < h_binstart[b + 1]; i += per) units.push_back(make_int2(b, i));
        }
    }

    int upload(const float* rest4) {
        size_t v4 = sizeof(float4) * m_n;
        bool ok = RX_HIP(hipMalloc(&d_x, v4)) && RX_HIP(hipMalloc(&d_v, v4)) && RX_HIP(hipMalloc(&d_xa, v4)) &&
                  RX_HIP(hipMalloc(&d_xb, v4)) && RX_HIP(hipMalloc(&d_rest, v4)) && RX_HIP(hipMalloc(&d_pre, v4 * 3)) &&
                  RX_HIP(hipMalloc(&d_tr, v4 * 3)) && RX_HIP(hipMalloc(&d_cs, sizeof(int) * (m_n + 1))) &&
                  RX_HIP(hipMalloc(&d_mem, sizeof(int) * std::max<size_t>(h_mem.size(), 1))) &&
                  RX_HIP(hipMalloc(&d_ps, sizeof(int) * (m_n + 1))) && RX_HIP(hipMalloc(&d_pc, sizeof(int) * std::max<size_t>(h_pc.size(), 1))) &&
                  RX_HIP(hipMalloc(&d_units, sizeof(int2) * (m_n + kNumBins))) && RX_HIP(hipMalloc(&d_tiles, sizeof(int) * kNumBins)) &&
                  RX_HIP(hipMalloc(&d_binend, sizeof(int) * kNumBins)) && RX_HIP(hipMalloc(&d_corder, sizeof(int) * m_n)) &&
                  RX_HIP(hipMalloc(&d_act, sizeof(int) * m_n));
        if (!ok) return 1;
        hipStream_t s = m_stream;
        std::vector<float4> pre(3 * (size_t)m_n);
        for (uint c = 0; c < m_n; ++c) {
            const float* o = &h_pre[(size_t)c * 12];
            pre[3 * c] = make_float4(o[0], o[1], o[2], o[3]);
            pre[3 * c + 1] = make_float4(o[4], o[5], o[6], o[7]);
            pre[3 * c + 2] = make_float4(o[8], o[9], o[10], o[11]);
        }
        RX_HIP(hipMemcpyAsync(d_rest, rest4, v4, hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_x, rest4, v4, hipMemcpyHostToDevice, s));
        rx_fill<float4>(d_v, make_float4(0, 0, 0, 0), m_n, s);
        RX_HIP(hipMemcpyAsync(d_pre, pre.data(), sizeof(float4) * pre.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_cs, h_cs.data(), sizeof(int) * h_cs.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_mem, h_mem.data(), sizeof(int) * h_mem.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_ps, h_ps.data(), sizeof(int) * h_ps.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_pc, h_pc.data(), sizeof(int) * h_pc.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_corder, h_corder.data(), sizeof(int) * m_n, hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_binend, h_binstart.data() + 1, sizeof(int) * kNumBins, hipMemcpyHostToDevice, s));
        rx_fill<int>(d_act, 1, m_n, s);
        RX_HIP(hipStreamSynchronize(s));
        return push_tiles();
    }

    // Melt mask from the host (1 = ice); NULL removes the mask.
    int set_active(const int* host) {
        if (host) {
            RX_HIP(hipMemcpyAsync(d_act, host, sizeof(int) * m_n, hipMemcpyHostToDevice, m_stream));
        } else {
            rx_fill<int>(d_act, 1, m_n, m_stream);
        }
        m_masked = host != nullptr;
        drop_graph();  // the step's launch parameters changed
        return RX_HIP(hipStreamSynchronize(m_stream)) ? 0 : 1;
    }
    // Device view for a caller that drives the object from its own kernels (the coupled solver): state arrays
    // in ice-particle order, the mask, the stream.  Turns the mask on.
    void device_view(float4** x, float4** v, int** act, hipStream_t* s) {
        if (!m_masked) {
            m_masked = true;
            drop_graph();
        }
        *x = d_x;
        *v = d_v;
        *act = d_act;
        *s = m_stream;
    }

    int push_tiles() {
        std::vector<int2> units;
        make_units(units);
        m_nunits = (int)units.size();
        hipStream_t s = m_stream;
        RX_HIP(hipMemcpyAsync(d_units, units.data(), sizeof(int2) * units.size(), hipMemcpyHostToDevice, s));
        RX_HIP(hipMemcpyAsync(d_tiles, m_tiles, sizeof(int) * kNumBins, hipMemcpyHostToDevice, s));
        drop_graph();
        return RX_HIP(hipStreamSynchronize(s)) ? 0 : 1;
    }

    SmK consts() const {
        SmK k;
        k.alpha = m_env.alpha;
        k.beta = m_env.beta;
        k.dt = m_env.dt;
        k.res = m_env.restitution;
        k.g = f3(m_env.gravity[0], m_env.gravity[1], m_env.gravity[2]);
        k.lo = f3(m_env.box_min[0], m_env.box_min[1], m_env.box_min[2]);
        k.hi = f3(m_env.box_max[0], m_env.box_max[1], m_env.box_max[2]);
        k.volcon = m_env.volume_conservation;
        k.n = m_n;
        k.masked = m_masked ? 1 : 0;
        return k;
    }

    int grid_blocks() const {
        int waves = (m_nunits + 0) ;
        int blocks = (waves + 3) / 4;
        return std::max(1, std::min(blocks, m_max_blocks));
    }

    void launch_clusters(hipStream_t s, const int2* units, int nunits, const float4* xs) {
        SmK k = consts();
        int blocks = std::max(1, std::min((nunits + 3) / 4, m_max_blocks));
        k_sm_clusters<<<blocks, 256, 0, s>>>(units, nunits, d_tiles, d_binend, d_corder, d_cs, d_mem, xs, d_rest, d_pre, d_tr, d_act, k);
    }

    // One step as a fixed launch sequence (no pointer swaps: graph capturable).
    void enqueue_step(bool timed) {
        hipStream_t s = m_stream;
        SmK k = consts();
        unsigned nb = rx_blocks(m_n, 256);
        k_sm_predict<<<nb, 256, 0, s>>>(d_x, d_v, d_xa, k);
        float4* cur = d_xa;
        float4* nxt = d_xb;
        for (int it = 0; it < m_env.iterations; ++it) {
            if (timed) RX_HIP(hipEventRecord(m_ev[0], s));
            launch_clusters(s, d_units, m_nunits, cur);
            if (timed) RX_HIP(hipEventRecord(m_ev[1], s));
            k_sm_average<<<nb, 256, 0, s>>>(d_ps, d_pc, d_tr, d_rest, cur, nxt, d_act, k);
            if (timed) RX_HIP(hipEventRecord(m_ev[2], s));
            std::swap(cur, nxt);
            if (timed) {
                RX_HIP(hipStreamSynchronize(s));
                float a = 0, b = 0;
                RX_HIP(hipEventElapsedTime(&a, m_ev[0], m_ev[1]));
                RX_HIP(hipEventElapsedTime(&b, m_ev[1], m_ev[2]));
                m_t_cluster += a;
                m_t_avg += b;
                m_t_n += 1;
            }
        }
        k_sm_velocity<<<nb, 256, 0, s>>>(d_x, d_v, cur, k);
    }

    int update(int steps) {
        for (int t = 0; t < steps; ++t) {
            if (!m_use_graph) {
                enqueue_step(false);
                continue;
            }
            if (!m_graph) {
                hipGraph_t g;
                if (!RX_HIP(hipStreamBeginCapture(m_stream, hipStreamCaptureModeThreadLocal))) return 1;
                enqueue_step(false);
                if (!RX_HIP(hipStreamEndCapture(m_stream, &g))) return 1;
                if (!RX_HIP(hipGraphInstantiate(&m_graph, g, nullptr, nullptr, 0))) return 1;
                RX_HIP(hipGraphDestroy(g));
            }
            if (!RX_HIP(hipGraphLaunch(m_graph, m_stream))) return 1;
        }
        return RX_HIP(hipGetLastError()) ? 0 : 1;
    }

    int kernel_times(float* cl, float* av) {
        ensure_events();
        m_t_cluster = m_t_avg = 0.0;
        m_t_n = 0;
        // timed eager step on a copy of the state
        std::vector<float4> x(m_n), v(m_n);
        RX_HIP(hipMemcpyAsync(x.data(), d_x, sizeof(float4) * m_n, hipMemcpyDeviceToHost, m_stream));
        RX_HIP(hipMemcpyAsync(v.data(), d_v, sizeof(float4) * m_n, hipMemcpyDeviceToHost, m_stream));
        RX_HIP(hipStreamSynchronize(m_stream));
        for (int r = 0; r < 3; ++r) enqueue_step(true);
        RX_HIP(hipMemcpyAsync(d_x, x.data(), sizeof(float4) * m_n, hipMemcpyHostToDevice, m_stream));
        RX_HIP(hipMemcpyAsync(d_v, v.data(), sizeof(float4) * m_n, hipMemcpyHostToDevice, m_stream));
        RX_HIP(hipStreamSynchronize(m_stream));
        if (cl) *cl = m_t_n ? (float)(m_t_cluster / m_t_n) : 0.0f;
        if (av) *av = m_t_n ? (float)(m_t_avg / m_t_n) : 0.0f;
        return 0;
    }

    // Times each candidate tile per bin (cluster kernel on that bin alone) and each persistent grid size.
    int autotune(int* out) {
        ensure_events();
        hipStream_t s = m_stream;
        k_sm_predict<<<rx_blocks(m_n, 256), 256, 0, s>>>(d_x, d_v, d_xa, consts());
        const int cand[4] = {8, 16, 32, 64};
        std::vector<int2> units;
        for (int b = 0; b < kNumBins; ++b) {
            if (h_binstart[b + 1] == h_binstart[b]) continue;
            float best = 1e30f;
            int best_t = m_tiles[b];
            for (int T : cand) {
                m_tiles[b] = T;
                RX_HIP(hipMemcpyAsync(d_tiles, m_tiles, sizeof(int) * kNumBins, hipMemcpyHostToDevice, s));
                make_units(units, b);
                RX_HIP(hipMemcpyAsync(d_units, units.data(), sizeof(int2) * units.size(), hipMemcpyHostToDevice, s));
                launch_clusters(s, d_units, (int)units.size(), d_xa);  // warm
                RX_HIP(hipEventRecord(m_ev[0], s));
                for (int r = 0; r < 5; ++r) launch_clusters(s, d_units, (int)units.size(), d_xa);
                RX_HIP(hipEventRecord(m_ev[1], s));
                RX_HIP(hipStreamSynchronize(s));
                float ms = 0;
                RX_HIP(hipEventElapsedTime(&ms, m_ev[0], m_ev[1]));
                if (ms < best) {
                    best = ms;
                    best_t = T;
                }
            }
            m_tiles[b] = best_t;
        }
        // persistent grid size
        push_tiles();
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) cus = prop.multiProcessorCount;
        const int mult[5] = {2, 4, 8, 16, 1 << 20};
        float best = 1e30f;
        int best_b = m_max_blocks;
        for (int m : mult) {
            m_max_blocks = m >= (1 << 20) ? (1 << 30) : cus * m;
            launch_clusters(s, d_units, m_nunits, d_xa);
            RX_HIP(hipEventRecord(m_ev[0], s));
            for (int r = 0; r < 5; ++r) launch_clusters(s, d_units, m_nunits, d_xa);
            RX_HIP(hipEventRecord(m_ev[1], s));
            RX_HIP(hipStreamSynchronize(s));
            float ms = 0;
            RX_HIP(hipEventElapsedTime(&ms, m_ev[0], m_ev[1]));
            if (ms < best) {
                best = ms;
                best_b = m_max_blocks;
            }
        }
        m_max_blocks = best_b;
        drop_graph();
        if (out)
            for (int b = 0; b < kNumBins; ++b) out[b] = m_tiles[b];
        return RX_HIP(hipGetLastError()) ? 0 : 1;
    }

    void ensure_events() {
        while (m_ev.size() < 3) {
            hipEvent_t e;
            RX_HIP(hipEventCreate(&e));
            m_ev.push_back(e);
        }
    }
    void drop_graph() {
        if (m_graph) {
            if (m_stream) RX_HIP(hipStreamSynchronize(m_stream));  // no launch of the graph is still pending
            RX_HIP(hipGraphExecDestroy(m_graph));
            m_graph = nullptr;
        }
    }

    rxIceEnv m_env{};
    uint m_n = 0;
    hipStream_t m_stream = nullptr;
    hipGraphExec_t m_graph = nullptr;
    bool m_use_graph = true;
    int m_tiles[kNumBins] = {8, 16, 32, 64, 64};
    int m_max_blocks = 1 << 30;
    int m_nunits = 0;
    std::vector<int> h_cs, h_mem, h_ps, h_pc, h_corder, h_binstart;
    std::vector<float> h_pre;
    float4 *d_x = nullptr, *d_v = nullptr, *d_xa = nullptr, *d_xb = nullptr, *d_rest = nullptr, *d_pre = nullptr, *d_tr = nullptr;
    int* d_act = nullptr;  // melt mask: 1 = the particle is ice
    bool m_masked = false;
    int *d_cs = nullptr, *d_mem = nullptr, *d_ps = nullptr, *d_pc = nullptr, *d_tiles = nullptr, *d_binend = nullptr, *d_corder = nullptr;
    int2* d_units = nullptr;
    std::vector<hipEvent_t> m_ev;
    double m_t_cluster = 0, m_t_avg = 0;
    long m_t_n = 0;
};

extern "C" {

void rxIceEnvDefault(rxIceEnv* e) {
    memset(e, 0, sizeof(*e));
    e->alpha = 1.0f;
    e->beta = 0.0f;
    e->dt = 0.01f;
    e->gravity[1] = -9.81f;  // ShapeMatching.cpp:46
    e->box_min[0] = e->box_min[1] = e->box_min[2] = -1.0f;
    e->box_max[0] = e->box_max[1] = e->box_max[2] = 1.0f;
    e->iterations = 1;
    e->volume_conservation = 1;
    e->restitution = 0.9f;  // ElasticObject_OP.cpp:580
    e->cluster_radius = 0.0f;
}

void* rxIceSM_Create(const rxIceEnv* env, const float* rest_pos4, int n) {
    if (!env || !rest_pos4 || n <= 0 || !(env->cluster_radius > 0.0f) || !(env->dt > 0.0f)) {
        rx_set_error(1, "rxIceSM_Create: bad arguments");
        return nullptr;
    }
    rxIceSM* p = new rxIceSM();
    if (p->init(*env, rest_pos4, n) != 0) {
        delete p;
        return nullptr;
    }
    return p;
}
void rxIceSM_Destroy(void* h) { delete (rxIceSM*)h; }
int rxIceSM_SetState(void* h, const float* pos4, const float* vel4) {
    if (!h || !pos4) return 1;
    rxIceSM* p = (rxIceSM*)h;
    RX_HIP(hipMemcpyAsync(p->d_x, pos4, sizeof(float4) * p->m_n, hipMemcpyHostToDevice, p->m_stream));
    if (vel4) RX_HIP(hipMemcpyAsync(p->d_v, vel4, sizeof(float4) * p->m_n, hipMemcpyHostToDevice, p->m_stream));
    else rx_fill<float4>(p->d_v, make_float4(0, 0, 0, 0), p->m_n, p->m_stream);
    return RX_HIP(hipStreamSynchronize(p->m_stream)) ? 0 : 1;
}
int rxIceSM_GetState(void* h, float* pos4, float* vel4) {
    if (!h) return 1;
    rxIceSM* p = (rxIceSM*)h;
    if (pos4) RX_HIP(hipMemcpyAsync(pos4, p->d_x, sizeof(float4) * p->m_n, hipMemcpyDeviceToHost, p->m_stream));
    if (vel4) RX_HIP(hipMemcpyAsync(vel4, p->d_v, sizeof(float4) * p->m_n, hipMemcpyDeviceToHost, p->m_stream));
    return RX_HIP(hipStreamSynchronize(p->m_stream)) ? 0 : 1;
}
int rxIceSM_Update(void* h, int steps) { return h ? ((rxIceSM*)h)->update(steps) : 1; }
int rxIceSM_Sync(void* h) { return h ? (RX_HIP(hipStreamSynchronize(((rxIceSM*)h)->m_stream)) ? 0 : 1) : 1; }
int rxIceSM_ClusterInfo(void* h, int* nclusters, long* nmembers, int* hist, int nhist) {
    if (!h) return 1;
    rxIceSM* p = (rxIceSM*)h;
    if (nclusters) *nclusters = (int)p->m_n;
    if (nmembers) *nmembers = (long)p->h_mem.size();
    if (hist) {
        for (int k = 0; k < nhist; ++k) hist[k] = 0;
        for (uint c = 0; c < p->m_n; ++c) {
            int s = p->h_cs[c + 1] - p->h_cs[c];
            if (s < nhist) hist[s]++;
        }
    }
    return 0;
}
int rxIceSM_GetClusters(void* h, int* start, int* members) {
    if (!h) return 1;
    rxIceSM* p = (rxIceSM*)h;
    if (start) memcpy(start, p->h_cs.data(), sizeof(int) * p->h_cs.size());
    if (members) memcpy(members, p->h_mem.data(), sizeof(int) * p->h_mem.size());
    return 0;
}
int rxIceSM_Autotune(void* h, int* out_tiles) { return h ? ((rxIceSM*)h)->autotune(out_tiles) : 1; }
int rxIceSM_SetTiles(void* h, const int* tiles) {
    if (!h || !tiles) return 1;
    rxIceSM* p = (rxIceSM*)h;
    for (int b = 0; b < kNumBins; ++b) {
        int t = tiles[b];
        if (t != 8 && t != 16 && t != 32 && t != 64) {
            rx_set_error(1, "rxIceSM_SetTiles: tiles must be 8, 16, 32 or 64");
            return 1;
        }
        p->m_tiles[b] = t;
    }
    return p->push_tiles();
}
int rxIceSM_UseGraph(void* h, int on) {
    if (!h) return 1;
    rxIceSM* p = (rxIceSM*)h;
    p->m_use_graph = on != 0;
    if (!on) p->drop_graph();
    return 0;
}
int rxIceSM_KernelTimes(void* h, float* cl, float* av) { return h ? ((rxIceSM*)h)->kernel_times(cl, av) : 1; }
int rxIceSM_SetActive(void* h, const int* active) { return h ? ((rxIceSM*)h)->set_active(active) : 1; }

}  // extern "C"

// Internal (not part of the C ABI): device view of an rxIceSM for the coupled solver.
bool rx_ice_device_view(void* h, float4** x, float4** v, int** act, hipStream_t* s) {
    if (!h) return false;
    ((rxIceSM*)h)->device_view(x, v, act, s);
    return true;
}

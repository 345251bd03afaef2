// ice_launch.hip -- the C++-linkage ice launch functions of Ice_SM.h:23-53 (LaunchShapeMatchingGPU,
// LaunchShapeMatchingIterationGPU) and IceObject.h:370-395 (LaunchCalcAverageGPU, LaunchInterPolationGPU).
#include "rx_common.h"
#include "../../include/rx_ice_funcs.h"

namespace {

__device__ inline void jacobi3d(double a[9], double v[9]) {
    // EigenJacobiMethod (ShapeMatching.h:242-343), n = 3, eps 1e-8, at most 100 rotations (double)
    for (int cnt = 0; cnt <= 100; ++cnt) {
        int i = -1, j = -1;
        double x = 0.0;
        for (int ia = 0; ia < 3; ++ia)
            for (int ja = 0; ja < 3; ++ja)
                if (ia != ja && fabs(a[ia * 3 + ja]) > x) {
                    i = ia;
                    j = ja;
                    x = fabs(a[ia * 3 + ja]);
                }
        if (i == -1) return;
        double aii = a[i * 3 + i], ajj = a[j * 3 + j], aij = a[i * 3 + j];
        double al = (aii - ajj) / 2.0, be = sqrt(al * al + aij * aij);
        double ct = sqrt((1.0 + fabs(al) / be) / 2.0);
        double st = (((aii - ajj) >= 0.0) ? 1.0 : -1.0) * aij / (2.0 * be * ct);
        int m = 3 - i - j;
        double aim = a[i * 3 + m], ajm = a[j * 3 + m];
        a[i * 3 + m] = a[m * 3 + i] = aim * ct + ajm * st;
        a[j * 3 + m] = a[m * 3 + j] = -aim * st + ajm * ct;
        a[i * 3 + i] = aii * ct * ct + 2.0 * aij * ct * st + ajj * st * st;
        a[j * 3 + j] = aii * st * st - 2.0 * aij * ct * st + ajj * ct * ct;
        a[i * 3 + j] = a[j * 3 + i] = 0.0;
        for (int r = 0; r < 3; ++r) {
            double vi = v[r * 3 + i], vj = v[r * 3 + j];
            v[r * 3 + i] = vi * ct + vj * st;
            v[r * 3 + j] = -vi * st + vj * ct;
        }
        double e = 2.0 * (fabs(a[1]) + fabs(a[2]) + fabs(a[5]));
        if (e < 1e-8) return;
    }
}

// PolarDecomposition (ShapeMatching.h:352-387): R = A (A^T A)^(-1/2) from the Jacobi eigenvectors of A^T A.
__device__ inline void polar_rotation(const double apq[9], double R[9]) {
    double ata[9], U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) ata[r * 3 + cc] = apq[r] * apq[cc] + apq[3 + r] * apq[3 + cc] + apq[6 + r] * apq[6 + cc];
    jacobi3d(ata, U);
    double l[3];
    for (int t = 0; t < 3; ++t) l[t] = ata[t * 4] <= 0.0 ? 0.0 : 1.0 / sqrt(ata[t * 4]);
    double S1[9];
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) S1[r * 3 + cc] = l[0] * U[r * 3] * U[cc * 3] + l[1] * U[r * 3 + 1] * U[cc * 3 + 1] + l[2] * U[r * 3 + 2] * U[cc * 3 + 2];
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) R[r * 3 + cc] = apq[r * 3] * S1[cc] + apq[r * 3 + 1] * S1[3 + cc] + apq[r * 3 + 2] * S1[6 + cc];
}

// OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:551-720), one thread per cluster with the
// reference's slot layout (ShapeMatching.cu:57-70 launched it as one thread per cluster too).
__global__ void k_update_cluster_sm(int nc, const float4* __restrict__ prtPos, const float4* __restrict__ prtVel,
                                    const float* __restrict__ orgPos, float* __restrict__ curPos, const float* __restrict__ orgCm,
                                    float* __restrict__ curCm, float* __restrict__ vel, const int* __restrict__ pIndxes,
                                    const int* __restrict__ indxSet, float dt, float3 g, float3 lo, float3 hi) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int s = indxSet[2 * c], e = indxSet[2 * c + 1];
    const double res = 0.9;
    double gg[3] = {g.x, g.y, g.z}, mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
    for (int i = s; i <= e; ++i) {
        int p = pIndxes[i];
        float4 x4 = prtPos[p], v4 = prtVel[p];
        double x[3] = {x4.x, x4.y, x4.z}, v[3] = {v4.x, v4.y, v4.z}, cp[3];
        for (int a = 0; a < 3; ++a) cp[a] = (float)(x[a] + (v[a] + gg[a] * dt) * dt);
        for (int a = 0; a < 3; ++a)
            if (cp[a] < mn[a] || cp[a] > mx[a]) {
                cp[a] = (float)(x[a] - (v[a] + gg[a] * dt) * dt * res);
                for (int o = 0; o < 3; ++o)
                    if (o != a) cp[o] = x[o];
            }
        for (int a = 0; a < 3; ++a) curPos[3 * i + a] = (float)fmin(fmax(cp[a], mn[a]), mx[a]);
    }
    int k = e - s + 1;
    if (k <= 1) return;  // ElasticObject_OP.cpp:611
    double cm[3] = {0, 0, 0};
    for (int i = s; i <= e; ++i)
        for (int a = 0; a < 3; ++a) cm[a] += curPos[3 * i + a];
    for (int a = 0; a < 3; ++a) cm[a] /= k;
    double apq[9] = {0};
    for (int i = s; i <= e; ++i) {
        double p[3], q[3];
        for (int a = 0; a < 3; ++a) {
            p[a] = curPos[3 * i + a] - cm[a];
            q[a] = orgPos[3 * i + a] - orgCm[3 * c + a];
        }
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) apq[r * 3 + cc] += p[r] * q[cc];
    }
    double R[9];
    polar_rotation(apq, R);
    for (int i = s; i <= e; ++i) {
        double q[3];
        for (int a = 0; a < 3; ++a) q[a] = orgPos[3 * i + a] - orgCm[3 * c + a];
        for (int a = 0; a < 3; ++a) curPos[3 * i + a] = (float)(R[a * 3] * q[0] + R[a * 3 + 1] * q[1] + R[a * 3 + 2] * q[2] + cm[a]);
    }
    for (int a = 0; a < 3; ++a) curCm[3 * c + a] = (float)cm[a];
    double idt = 1.0 / dt;
    for (int i = s; i <= e; ++i) {
        int p = pIndxes[i];
        float4 x4 = prtPos[p];
        double x[3] = {x4.x, x4.y, x4.z};
        for (int a = 0; a < 3; ++a) vel[3 * i + a] = (float)((curPos[3 * i + a] - x[a]) * idt);
    }
}

// OrientedPrtObj::UpdateCluster_SM_Itr (ElasticObject_OP.cpp:722-887), one thread per cluster: the members take
// the averaged solid positions of the last pass (sldPos, 3 floats per particle), the cluster is matched rigidly
// (Apq's third column negated when det Apq < 0 and the cluster has at least 10 members), and the goal positions,
// clamped to the walls, become curPos.  The velocities are left alone (the reference's iteration pass does not
// write them).  A one-member cluster keeps the copied position.
__global__ void k_update_cluster_sm_itr(int nc, const float* __restrict__ sldPos, const float* __restrict__ orgPos,
                                        float* __restrict__ curPos, const float* __restrict__ orgCm, float* __restrict__ curCm,
                                        const int* __restrict__ pIndxes, const int* __restrict__ indxSet, float3 lo, float3 hi) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int s = indxSet[2 * c], e = indxSet[2 * c + 1];
    for (int i = s; i <= e; ++i) {
        int p = pIndxes[i];
        for (int a = 0; a < 3; ++a) curPos[3 * i + a] = sldPos[3 * p + a];
    }
    int k = e - s + 1;
    if (k <= 1) return;
    double cm[3] = {0, 0, 0};
    for (int i = s; i <= e; ++i)
        for (int a = 0; a < 3; ++a) cm[a] += curPos[3 * i + a];
    for (int a = 0; a < 3; ++a) cm[a] /= k;
    double apq[9] = {0};
    for (int i = s; i <= e; ++i) {
        double p[3], q[3];
        for (int a = 0; a < 3; ++a) {
            p[a] = curPos[3 * i + a] - cm[a];
            q[a] = orgPos[3 * i + a] - orgCm[3 * c + a];
        }
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) apq[r * 3 + cc] += p[r] * q[cc];
    }
    const double det = apq[0] * (apq[4] * apq[8] - apq[5] * apq[7]) - apq[1] * (apq[3] * apq[8] - apq[5] * apq[6]) +
                       apq[2] * (apq[3] * apq[7] - apq[4] * apq[6]);
    if (k >= 10 && det < 0.0)
        for (int r = 0; r < 3; ++r) apq[r * 3 + 2] = -apq[r * 3 + 2];
    double R[9];
    polar_rotation(apq, R);
    const double mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {hi.x, hi.y, hi.z};
    for (int i = s; i <= e; ++i) {
        double q[3];
        for (int a = 0; a < 3; ++a) q[a] = orgPos[3 * i + a] - orgCm[3 * c + a];
        for (int a = 0; a < 3; ++a) {
            double gp = (double)(float)(R[a * 3] * q[0] + R[a * 3 + 1] * q[1] + R[a * 3 + 2] * q[2] + cm[a]);
            curPos[3 * i + a] = (float)fmin(fmax(gp, mn[a]), mx[a]);
        }
    }
    for (int a = 0; a < 3; ++a) curCm[3 * c + a] = (float)cm[a];
}

// ConvoNormal::StepConvolution (Ice_Convolution_Normal.cpp:32-67) in the PtoC layout of IceStructure.cu:63-126,
// one thread per particle: the mean member position and velocity over the clusters that hold the particle.
// Entry j of particle p (j < min(PtoCIndx[p], PtoCMax)) names cluster PtoC[base + j] and member number
// PtoC[base + PtoCMax + j], base = PtoCMax * PtoCParamSize * p; -1 marks an empty entry.  The sums start at zero
// and a particle in no cluster keeps its solid position and velocity.
__global__ void k_calc_average(int n, float* __restrict__ sldPos, float* __restrict__ sldVel, const float* __restrict__ smPos,
                               const float* __restrict__ smVel, const int* __restrict__ indxSet, const int* __restrict__ PtoCIndx,
                               const int* __restrict__ PtoC, int PtoCMax, int PtoCParamSize) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long base = (long)PtoCMax * PtoCParamSize * p;
    const int cnt = min(PtoCIndx[p], PtoCMax);
    double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    int k = 0;
    for (int j = 0; j < cnt; ++j) {
        const int c = PtoC[base + j], o = PtoC[base + PtoCMax + j];
        if (c == -1 || o == -1) continue;
        const long slot = (long)indxSet[2 * c] + o;
        for (int a = 0; a < 3; ++a) {
            sp[a] += smPos[3 * slot + a];
            sv[a] += smVel[3 * slot + a];
        }
        ++k;
    }
    if (k == 0) return;
    for (int a = 0; a < 3; ++a) {
        sldPos[3 * p + a] = (float)(sp[a] / k);
        sldVel[3 * p + a] = (float)(sv[a] / k);
    }
}

// IceInterPolation.cu:33-49 (LinerInterPolation)
__global__ void k_interpolation(int n, const float* __restrict__ sp, const float* __restrict__ sv, float* __restrict__ fp, float* __restrict__ fv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < 3; ++a) {
        fp[4 * i + a] = sp[3 * i + a];
        fv[4 * i + a] = sv[3 * i + a];
    }
}

}  // namespace

void LaunchShapeMatchingGPU(int prtNum, float* prtPos, float* prtVel, float* orgPos, float* curPos, float* orgCm, float* curCm, float* vel,
                            int* pIndxes, int* d_IndxSet, float dt) {
    if (prtNum <= 0) return;
    if (!rx_params_set()) {
        rx_set_error(3, "LaunchShapeMatchingGPU: CuSetParameters must be called first (gravity and walls)");
        return;
    }
    const rxSimParams& p = rx_params();
    hipStream_t s = rx_stream();
    k_update_cluster_sm<<<rx_blocks(prtNum, 64), 64, 0, s>>>(prtNum, (const float4*)prtPos, (const float4*)prtVel, orgPos, curPos, orgCm,
                                                             curCm, vel, pIndxes, d_IndxSet, dt, p.Gravity, p.BoundaryMin, p.BoundaryMax);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void LaunchInterPolationGPU(int prtNum, float* sldPrtPos, float* sldPrtVel, float* sphPrtPos, float* sphPrtVel) {
    if (prtNum <= 0) return;
    hipStream_t s = rx_stream();
    k_interpolation<<<rx_blocks(prtNum, 256), 256, 0, s>>>(prtNum, sldPrtPos, sldPrtVel, sphPrtPos, sphPrtVel);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void LaunchShapeMatchingIterationGPU(int prtNum, float* prtPos, float* prtVel, float* sldPos, float* sldVel, float* orgPos, float* curPos,
                                     float* orgCm, float* curCm, float* vel, int* pIndxes, int* d_IndxSet, float dt) {
    (void)prtPos, (void)prtVel, (void)sldVel, (void)vel, (void)dt;  // the iteration pass reads only the solid positions
    if (prtNum <= 0) return;
    if (!rx_params_set()) {
        rx_set_error(3, "LaunchShapeMatchingIterationGPU: CuSetParameters must be called first (walls)");
        return;
    }
    const rxSimParams& p = rx_params();
    hipStream_t s = rx_stream();
    k_update_cluster_sm_itr<<<rx_blocks(prtNum, 64), 64, 0, s>>>(prtNum, sldPos, orgPos, curPos, orgCm, curCm, pIndxes, d_IndxSet,
                                                                 p.BoundaryMin, p.BoundaryMax);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void LaunchCalcAverageGPU(int prtNum, float* sldPrtPos, float* sldPrtVel, float* sphPrtPos, float* sphPrtVel, float* smPrtPos, float* smPrtVel,
                          int* smIndxSet, int* PtoCIndx, int* PtoC, int PNumMax, int PtoCMax, int PtoCParamSize) {
    (void)sphPrtPos, (void)sphPrtVel, (void)PNumMax;  // unused by the averaging, as in the reference kernel
    if (prtNum <= 0) return;
    if (PtoCMax <= 0 || PtoCParamSize < 2) {
        rx_set_error(1, "LaunchCalcAverageGPU: PtoCMax must be positive and PtoCParamSize at least 2");
        return;
    }
    hipStream_t s = rx_stream();
    k_calc_average<<<rx_blocks(prtNum, 256), 256, 0, s>>>(prtNum, sldPrtPos, sldPrtVel, smPrtPos, smPrtVel, smIndxSet, PtoCIndx, PtoC,
                                                          PtoCMax, PtoCParamSize);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

// aniso.hip -- anisotropic kernels for surface extraction (Yu & Turk 2013, "Reconstructing surfaces of
// particle-based fluids using anisotropic kernels"), behind the reference's entry points (rx_cu_funcs.cuh:
// 138-144, rx_sph.cu:1126-1276, rx_aniso_kernel.cu:20-718):
//
//   CuSphCalUpdatedPosition   weighted mean x_w of the neighbours within h (w = (1 - r/h)^3) and the smoothed
//                             kernel centre (1 - lambda) x + lambda x_w
//   CuSphCalCovarianceMatrix  C = sum w (x_j - x_w)(x_j - x_w)^T / sum w, neighbour count in PosW.w
//   CuSphSVDecomposition      C = R diag(sigma) R^T, sigma descending, clamped and scaled
//   CuSphCalTransformMatrix   G = R diag(1/sigma) R^T divided by its largest diagonal entry
//   CuSphGridDensityAniso     lattice density sum_j det(G_j) m W_poly6(|G_j (x - x_j)|, h) over the cells within 3h
//
// Summation order follows the reference: cells z, y, x over the clamped cell box, each cell's terms summed on
// their own and the cell sum added to the total (the covariance entries are summed straight into the total,
// as there).  Every multiply, add and divide rounds on its own (no contraction); with correctly rounded division
// (IEEE, or div_h below) and square root the results are bit-identical to the CPU restatement.
//
// The 3x3 decomposition is a cyclic Jacobi iteration in FP64 on the symmetric C (the reference runs a general
// SVD; for a symmetric positive semi-definite C the singular values are the eigenvalues and U = V).  The
// result is ordered and signed as the reference orders and signs its SVD: singular values |lambda| descending,
// ties in their original order; column k is negated when more than 3 of the 6 entries of u_k and v_k are
// negative (v_k = sign(lambda_k) u_k).
#include "rx_common.h"
#include "../../include/rx_cu_mesh.h"

namespace {

constexpr int kTpb = 256;

struct AnisoK {
    float3 origin, cellw;
    int gx, gy, gz;
    float h;       // EffectiveRadius (lattice density)
    float mass, wpoly6;
};

AnisoK aniso_consts() {
    const rxSimParams& p = rx_params();
    AnisoK k;
    k.origin = p.WorldOrigin;
    k.cellw = p.CellWidth;
    k.gx = (int)p.GridSize.x;
    k.gy = (int)p.GridSize.y;
    k.gz = (int)p.GridSize.z;
    k.h = p.EffectiveRadius;
    k.mass = p.Mass;
    k.wpoly6 = p.Wpoly6;
    return k;
}

#pragma clang fp contract(off)

// rx_cu_common.cu:120-133 (calcGridPos): floor((p - origin) / cell), clamped into the grid
__device__ inline int3 cell_of(float3 p, const AnisoK& k) {
    int3 g;
    g.x = min(max((int)floorf((p.x - k.origin.x) / k.cellw.x), 0), k.gx - 1);
    g.y = min(max((int)floorf((p.y - k.origin.y) / k.cellw.y), 0), k.gy - 1);
    g.z = min(max((int)floorf((p.z - k.origin.z) / k.cellw.z), 0), k.gz - 1);
    return g;
}

// Squared distance from p to cell (x, y, z), or +inf-like "never skip" (-1) for cells on the grid's faces (the
// clamped hash puts particles from outside the grid there).  An interior cell holds only particles with
// floor((q - origin) / cell) equal to its index, computed in FP32, so they lie in its nominal box widened by
// the rounding of that formula; the widening used is 1e-3 of a cell plus 1e-6 of the grid's extent.
__device__ inline float cell_dist2(float3 p, int x, int y, int z, const AnisoK& k) {
    if (x <= 0 || y <= 0 || z <= 0 || x >= k.gx - 1 || y >= k.gy - 1 || z >= k.gz - 1) return -1.0f;
    auto axis = [](float q, int c, float o, float w, int n) {
        const float slack = 1e-3f * w + 1e-6f * (fabsf(o) + (float)n * w);
        const float lo = o + (float)c * w - slack, hi = o + (float)(c + 1) * w + slack;
        const float d = fmaxf(fmaxf(lo - q, q - hi), 0.0f);
        return d * d;
    };
    return axis(p.x, x, k.origin.x, k.cellw.x, k.gx) + axis(p.y, y, k.origin.y, k.cellw.y, k.gy) + axis(p.z, z, k.origin.z, k.cellw.z, k.gz);
}

// Calls body(j) for the slots of each cell of the box [lo, hi], cells in z, y, x order; cell_done() after
// each non-empty cell.  far(cell, d2) may skip a cell whose every particle the body would reject (d2 = the
// cell's squared distance from p, negative when unknown): a skipped cell adds nothing, as in the reference
// where its terms are all rejected and its zero sum is added.
// Within a cell, fetch(j) loads what body(j, data) needs; four fetches are issued before their bodies run (in
// slot order), so a wavefront keeps several gathers in flight.
template <class Fetch, class Body, class Done, class Far>
__device__ inline void for_cells(float3 p, int3 lo, int3 hi, const AnisoK& k, const uint* __restrict__ cs, const uint* __restrict__ ce,
                                 Fetch&& fetch, Body&& body, Done&& cell_done, Far&& far) {
    for (int z = lo.z; z <= hi.z; ++z)
        for (int y = lo.y; y <= hi.y; ++y)
            for (int x = lo.x; x <= hi.x; ++x) {
                const uint h = (uint)((z * k.gy + y) * k.gx + x);
                if (far(h, cell_dist2(p, x, y, z, k))) continue;
                const uint s = cs[h];
                if (s == kEmpty) continue;
                const uint e = ce[h];
                uint j = s;
                for (; j + 4 <= e; j += 4) {
                    const auto a = fetch(j), b = fetch(j + 1), c = fetch(j + 2), d = fetch(j + 3);
                    body(j, a);
                    body(j + 1, b);
                    body(j + 2, c);
                    body(j + 3, d);
                }
                for (; j < e; ++j) body(j, fetch(j));
                cell_done();
            }
}

// r2 above this has sqrtf(r2) >= h: the square root can be skipped for candidates beyond it
__device__ inline float sqrt_skip2(float h) { return h * h * 1.000001f; }

// r / h correctly rounded, from inv_h = 1 / h (correctly rounded) and one correction with the exact remainder
// (Markstein, "Computation of elementary functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1),
// 1990): q0 = r inv_h is faithful, e = r - q0 h is exact in one fma, and q0 + e inv_h rounds to r / h.  The
// remainder must not underflow, which holds for r = 0 and for every r = sqrtf(r2) of a float r2 > 0 (r >= 2^-75).
__device__ inline float div_h(float r, float h, float inv_h) {
    const float q0 = r * inv_h;
    return fmaf(fmaf(-q0, h, r), inv_h, q0);
}

// rx_aniso_kernel.cu:33-63 (calWeighedAvgPositionCell), summed over the cells of p +- h
__device__ inline float4 weighted_mean_sum(float3 p, float h, const float4* __restrict__ sp, const uint* __restrict__ cs,
                                           const uint* __restrict__ ce, const AnisoK& k) {
    const int3 lo = cell_of(make_float3(p.x - h, p.y - h, p.z - h), k);
    const int3 hi = cell_of(make_float3(p.x + h, p.y + h, p.z + h), k);
    const float reach2 = h * h * 1.0001f;  // a cell farther than h (with margin) has no particle with r < h
    const float far2 = sqrt_skip2(h);
    const float inv_h = 1.0f / h;
    float4 tot = make_float4(0.0f, 0.0f, 0.0f, 0.0f), part = tot;
    for_cells(
        p, lo, hi, k, cs, ce, [&](uint j) { return sp[j]; },
        [&](uint, const float4& q) {
            const float3 d = make_float3(p.x - q.x, p.y - q.y, p.z - q.z);
            const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
            if (r2 > far2) return;
            const float r = sqrtf(r2);
            if (r < h) {
                const float a = 1.0f - div_h(r, h, inv_h);
                const float w = a * a * a;
                part.x = part.x + q.x * w;
                part.y = part.y + q.y * w;
                part.z = part.z + q.z * w;
                part.w = part.w + w;
            }
        },
        [&]() {
            tot.x = tot.x + part.x;
            tot.y = tot.y + part.y;
            tot.z = tot.z + part.z;
            tot.w = tot.w + part.w;
            part = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        },
        [&](uint, float d2) { return d2 > reach2; });
    return tot;
}

// rx_aniso_kernel.cu:72-108 (sphCalUpdatedPosition)
__global__ __launch_bounds__(kTpb) void k_aniso_smooth(float4* __restrict__ up, float4* __restrict__ wpos, float lambda, float h,
                                                       rxParticleCell c, AnisoK k) {
    const uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    const float4 p4 = c.dSortedPos[i];
    const float3 p = make_float3(p4.x, p4.y, p4.z);
    float4 s = weighted_mean_sum(p, h, c.dSortedPos, c.dCellStart, c.dCellEnd, k);
    const float3 m = make_float3(s.x / s.w, s.y / s.w, s.z / s.w);
    const uint o = c.dSortedIndex[i];
    wpos[o] = make_float4(m.x, m.y, m.z, 0.0f);
    const float a = (float)(1.0 - (double)lambda);
    up[o] = make_float4(a * p.x + lambda * m.x, a * p.y + lambda * m.y, a * p.z + lambda * m.z, 0.0f);
}

// rx_aniso_kernel.cu:118-229 (calCovarianceMatrixCell, sphCalCovarianceMatrix)
__global__ __launch_bounds__(kTpb) void k_aniso_cov(float4* __restrict__ posw, float* __restrict__ cmat, float h, rxParticleCell c,
                                                    AnisoK k) {
    const uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    const float4 p4 = c.dSortedPos[i];
    const float3 p = make_float3(p4.x, p4.y, p4.z);
    float4 s = weighted_mean_sum(p, h, c.dSortedPos, c.dCellStart, c.dCellEnd, k);
    const float3 m = make_float3(s.x / s.w, s.y / s.w, s.z / s.w);
    const int3 lo = cell_of(make_float3(p.x - h, p.y - h, p.z - h), k);
    const int3 hi = cell_of(make_float3(p.x + h, p.y + h, p.z + h), k);
    const float reach2 = h * h * 1.0001f;
    const float far2 = sqrt_skip2(h);
    const float inv_h = 1.0f / h;
    float cv[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float wsum = 0.0f, cnt = 0.0f, pw = 0.0f, pc = 0.0f;
    for_cells(
        p, lo, hi, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint, const float4& q) {
            const float3 dd = make_float3(p.x - q.x, p.y - q.y, p.z - q.z);
            const float r2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
            if (r2 > far2) return;
            const float r = sqrtf(r2);
            if (r < h) {
                const float a = 1.0f - div_h(r, h, inv_h);
                const float w = a * a * a;
                const float d[3] = {q.x - m.x, q.y - m.y, q.z - m.z};
#pragma unroll
                for (int r3 = 0; r3 < 3; ++r3)
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) cv[3 * r3 + c3] = cv[3 * r3 + c3] + w * d[r3] * d[c3];
                pw = pw + w;
                pc = (float)((double)pc + 1.0);
            }
        },
        [&]() {
            wsum = wsum + pw;
            cnt = cnt + pc;
            pw = 0.0f;
            pc = 0.0f;
        },
        [&](uint, float d2) { return d2 > reach2; });
    const uint o = c.dSortedIndex[i];
#pragma unroll
    for (int t = 0; t < 9; ++t) cmat[9 * (size_t)o + t] = cv[t] / wsum;
    posw[o] = make_float4(m.x, m.y, m.z, cnt);
}

// Symmetric 3x3 eigen-decomposition, cyclic Jacobi (Golub & Van Loan, Matrix Computations, 8.5.2) in FP64:
// a = V diag(a_kk) V^T on return.  Sweeps stop once the off-diagonal magnitude is at most 1e-15 of the
// diagonal's, or after 16 sweeps.
__device__ inline void jacobi_rotate(double a[3][3], double v[3][3], int p, int q) {
    const double apq = a[p][q];
    if (apq == 0.0) return;
    const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double cs = 1.0 / sqrt(t * t + 1.0);
    const double sn = t * cs;
    const double tau = sn / (1.0 + cs);
    a[p][p] = a[p][p] - t * apq;
    a[q][q] = a[q][q] + t * apq;
    a[p][q] = 0.0;
    a[q][p] = 0.0;
    const int r = 3 - p - q;
    const double arp = a[r][p], arq = a[r][q];
    a[r][p] = arp - sn * (arq + tau * arp);
    a[p][r] = a[r][p];
    a[r][q] = arq + sn * (arp - tau * arq);
    a[q][r] = a[r][q];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double vp = v[m][p], vq = v[m][q];
        v[m][p] = vp - sn * (vq + tau * vp);
        v[m][q] = vq + sn * (vp - tau * vq);
    }
}
__device__ inline void sym_eigen3(double a[3][3], double v[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[r][c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double dg = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off == 0.0 || off <= 1e-15 * dg) break;
        jacobi_rotate(a, v, 0, 1);
        jacobi_rotate(a, v, 0, 2);
        jacobi_rotate(a, v, 1, 2);
    }
}

// rx_aniso_kernel.cu:490-534 (sphSVDecomposition): neighbour count n > 10: sigma_2,3 >= sigma_1 / 4, then all
// times 1400; otherwise sigma = (0.5, 0.5, 0.5).  R holds u_k in column k.
__global__ __launch_bounds__(kTpb) void k_aniso_svd(const float* __restrict__ cmat, const float4* __restrict__ posw,
                                                    float* __restrict__ eigen, float* __restrict__ rmat, uint n) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    const int nn = (int)posw[i].w;
    double a[3][3], u[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[r][c] = (double)cmat[9 * (size_t)i + 3 * r + c];
    // the decomposition works on the symmetric part (C is symmetric up to the rounding of its two halves)
    a[0][1] = a[1][0] = 0.5 * (a[0][1] + a[1][0]);
    a[0][2] = a[2][0] = 0.5 * (a[0][2] + a[2][0]);
    a[1][2] = a[2][1] = 0.5 * (a[1][2] + a[2][1]);
    sym_eigen3(a, u);
    float w[3], uf[3][3], vs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] = (float)fabs(a[k][k]);
        vs[k] = a[k][k] < 0.0 ? -1.0f : 1.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) uf[r][k] = (float)u[r][k];
    }
    // descending insertion sort, equal values keep their order
    for (int k = 1; k < 3; ++k) {
        for (int j = k; j > 0 && w[j - 1] < w[j]; --j) {
            float t = w[j - 1];
            w[j - 1] = w[j];
            w[j] = t;
            t = vs[j - 1];
            vs[j - 1] = vs[j];
            vs[j] = t;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                t = uf[r][j - 1];
                uf[r][j - 1] = uf[r][j];
                uf[r][j] = t;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int neg = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r) neg += (uf[r][k] < 0.0f) + (vs[k] * uf[r][k] < 0.0f);
        if (neg > 3)
#pragma unroll
            for (int r = 0; r < 3; ++r) uf[r][k] = -uf[r][k];
    }
    float s0 = w[0], s1 = w[1], s2 = w[2];
    if (nn > 10) {
        const float lo = s0 / 4.0f;
        s1 = s1 >= lo ? s1 : lo;
        s2 = s2 >= lo ? s2 : lo;
        s0 = s0 * 1400.0f;
        s1 = s1 * 1400.0f;
        s2 = s2 * 1400.0f;
    } else {
        s0 = s1 = s2 = 0.5f;
    }
    eigen[3 * (size_t)i + 0] = s0;
    eigen[3 * (size_t)i + 1] = s1;
    eigen[3 * (size_t)i + 2] = s2;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) rmat[9 * (size_t)i + 3 * r + c] = uf[r][c];
}

// rx_aniso_kernel.cu:583-616 (sphCalTransformMatrix)
__global__ __launch_bounds__(kTpb) void k_aniso_transform(const float* __restrict__ eigen, const float* __restrict__ rmat,
                                                          float* __restrict__ gmat, uint n) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= n) return;
    const float sx = eigen[3 * (size_t)i], sy = eigen[3 * (size_t)i + 1], sz = eigen[3 * (size_t)i + 2];
    float R[3][3], G[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[r][c] = rmat[9 * (size_t)i + 3 * r + c];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) G[j][c] = R[j][0] * R[c][0] / sx + R[j][1] * R[c][1] / sy + R[j][2] * R[c][2] / sz;
    float md = -100000000.0f;
    if (G[0][0] > md) md = G[0][0];
    if (G[1][1] > md) md = G[1][1];
    if (G[2][2] > md) md = G[2][2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) gmat[9 * (size_t)i + 3 * r + c] = G[r][c] / md;
}

// rx_aniso_kernel.cu:537-547 (calDeterminant), the reference's term order
__device__ inline float det3(const float* g) {
    return g[0] * g[4] * g[8] - g[0] * g[7] * g[5] + g[3] * g[7] * g[2] - g[3] * g[1] * g[8] + g[6] * g[1] * g[5] -
           g[6] * g[4] * g[2];
}

// Per sorted slot j: rec[j] = (x_j, mu2_j) with mu2_j a lower bound of the squared smallest singular value of
// G_j, and G_j's rows with det(G_j) in row 0's w.  Every r has |G_j r| >= sigma_min |r|, so a candidate with
// |r|^2 mu2_j > (1.01 h)^2 is outside the support |G_j r| <= h also after the FP32 rounding of G_j r (its
// relative error is below 1e-5 while sigma_min is at least 1e-3 of |G_j|; otherwise mu2_j = 0 and no candidate
// of j is rejected early).  sigma_min^2 is the smallest eigenvalue of G^T G (FP64 Jacobi).
__global__ __launch_bounds__(kTpb) void k_aniso_prep(const float* __restrict__ gmat, rxParticleCell c, float4* __restrict__ rec,
                                                     float4* __restrict__ gs) {
    const uint j = blockIdx.x * kTpb + threadIdx.x;
    if (j >= c.uNumParticles) return;
    const float* g = gmat + 9 * (size_t)c.dSortedIndex[j];
    float gl[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) gl[t] = g[t];
    double a[3][3], v[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            a[r][q] = (double)gl[r] * gl[q] + (double)gl[3 + r] * gl[3 + q] + (double)gl[6 + r] * gl[6 + q];
    const double tr = a[0][0] + a[1][1] + a[2][2];
    sym_eigen3(a, v);
    const double lmin = fmin(fmin(a[0][0], a[1][1]), a[2][2]);
    const float mu2 = (lmin > 1e-6 * tr) ? (float)(lmin * (1.0 - 1e-6)) : 0.0f;
    const float4 q = c.dSortedPos[j];
    rec[j] = make_float4(q.x, q.y, q.z, mu2);
    gs[3 * (size_t)j + 0] = make_float4(gl[0], gl[1], gl[2], det3(gl));
    gs[3 * (size_t)j + 1] = make_float4(gl[3], gl[4], gl[5], 0.0f);
    gs[3 * (size_t)j + 2] = make_float4(gl[6], gl[7], gl[8], 0.0f);
}

// Per cell: the smallest mu2 of its particles (+inf for an empty cell).
__global__ __launch_bounds__(kTpb) void k_aniso_cell_mu(const float4* __restrict__ rec, rxParticleCell c, float* __restrict__ cmu) {
    const uint h = blockIdx.x * kTpb + threadIdx.x;
    if (h >= c.uNumCells) return;
    const uint s = c.dCellStart[h];
    float m = __builtin_inff();
    if (s != kEmpty) {
        const uint e = c.dCellEnd[h];
        for (uint j = s; j < e; ++j) m = fminf(m, rec[j].w);
    }
    cmu[h] = m;
}

// rx_aniso_kernel.cu:625-715 (calDensityAnisoCellG, sphCalDensityAnisoInGrid).  Cells whose nearest point is
// beyond the support of all their particles (squared distance x the cell's smallest mu2 above (1.01 h)^2) are
// skipped, and so are candidates with |r|^2 mu2_j above it: neither can pass r <= h, so the sums are the
// reference's term for term.
__global__ __launch_bounds__(kTpb) void k_aniso_grid_density(float* __restrict__ out, const float4* __restrict__ rec,
                                                             const float4* __restrict__ gs, const float* __restrict__ cmu,
                                                             rxParticleCell c, AnisoK k, uint3 gnum, float3 gmin, float3 glen) {
    const size_t idx = (size_t)blockIdx.x * kTpb + threadIdx.x;
    const size_t total = (size_t)gnum.x * gnum.y * gnum.z;
    if (idx >= total) return;
    const uint gx = (uint)(idx % gnum.x), gy = (uint)((idx / gnum.x) % gnum.y), gz = (uint)(idx / ((size_t)gnum.x * gnum.y));
    const float3 p = make_float3(gmin.x + (float)gx * glen.x, gmin.y + (float)gy * glen.y, gmin.z + (float)gz * glen.z);
    const float h = k.h, h3 = h * 3.0f;
    const float reach2 = (h * 1.01f) * (h * 1.01f);
    const int3 lo = cell_of(make_float3(p.x - h3, p.y - h3, p.z - h3), k);
    const int3 hi = cell_of(make_float3(p.x + h3, p.y + h3, p.z + h3), k);
    float P = 0.0f, part = 0.0f;
    const float far2 = sqrt_skip2(h);
    for_cells(
        p, lo, hi, k, c.dCellStart, c.dCellEnd, [&](uint j) { return rec[j]; },
        [&](uint j, const float4& q) {
            const float3 rij = make_float3(p.x - q.x, p.y - q.y, p.z - q.z);
            if ((rij.x * rij.x + rij.y * rij.y + rij.z * rij.z) * q.w > reach2) return;
            const float4 g0 = gs[3 * (size_t)j], g1 = gs[3 * (size_t)j + 1], g2 = gs[3 * (size_t)j + 2];
            const float3 rg = make_float3(g0.x * rij.x + g0.y * rij.y + g0.z * rij.z, g1.x * rij.x + g1.y * rij.y + g1.z * rij.z,
                                          g2.x * rij.x + g2.y * rij.y + g2.z * rij.z);
            const float rg2 = rg.x * rg.x + rg.y * rg.y + rg.z * rg.z;
            if (rg2 > far2) return;
            const float r = sqrtf(rg2);
            if (r <= h) {
                const float qq = h * h - r * r;
                part = part + g0.w * k.mass * k.wpoly6 * qq * qq * qq;
            }
        },
        [&]() {
            P = P + part;
            part = 0.0f;
        },
        [&](uint cell, float d2) { return d2 >= 0.0f && d2 * cmu[cell] > reach2; });
    out[gx + gy * gnum.x + (size_t)gz * gnum.x * gnum.y] = P;
}

#pragma clang fp contract(on)

bool aniso_ready(const char* fn, const rxParticleCell& c, bool need_cells) {
    if (!rx_params_set()) {
        rx_set_error(3, "%s: CuSetParameters must be called first", fn);
        return false;
    }
    const rxSimParams& p = rx_params();
    if (need_cells && c.uNumParticles &&
        (!c.dSortedPos || !c.dSortedIndex || !c.dCellStart || !c.dCellEnd ||
         (size_t)p.GridSize.x * p.GridSize.y * p.GridSize.z != (size_t)c.uNumCells)) {
        rx_set_error(2, "%s: cell data does not match the grid of CuSetParameters (%u cells)", fn, c.uNumCells);
        return false;
    }
    return true;
}

void finish() {
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(rx_stream()));
}

}  // namespace

extern "C" {

void CuSphCalUpdatedPosition(float* dUpPos, float* dPosW, float lambda, float h, rxParticleCell cell) {
    if (!aniso_ready("CuSphCalUpdatedPosition", cell, true) || !cell.uNumParticles) return;
    k_aniso_smooth<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>((float4*)dUpPos, (float4*)dPosW, lambda, h, cell,
                                                                                   aniso_consts());
    finish();
}

void CuSphCalCovarianceMatrix(float* dPosW, float* dCMat, float h, rxParticleCell cell) {
    if (!aniso_ready("CuSphCalCovarianceMatrix", cell, true) || !cell.uNumParticles) return;
    k_aniso_cov<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>((float4*)dPosW, dCMat, h, cell, aniso_consts());
    finish();
}

void CuSphSVDecomposition(float* dC, float* dPosW, float* dEigen, float* dR, uint numParticles) {
    if (!numParticles) return;
    k_aniso_svd<<<rx_blocks(numParticles, kTpb), kTpb, 0, rx_stream()>>>(dC, (const float4*)dPosW, dEigen, dR, numParticles);
    finish();
}

void CuSphCalTransformMatrix(float* dEigen, float* dR, float* dG, uint numParticles) {
    if (!numParticles) return;
    k_aniso_transform<<<rx_blocks(numParticles, kTpb), kTpb, 0, rx_stream()>>>(dEigen, dR, dG, numParticles);
    finish();
}

void CuSphGridDensityAniso(float* dGridD, float* dG, float Emax, rxParticleCell cell, int nx, int ny, int nz, float x0, float y0,
                           float z0, float dx, float dy, float dz) {
    const char* fn = "CuSphGridDensityAniso";
    if (!aniso_ready(fn, cell, true)) return;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    const size_t total = (size_t)nx * ny * nz;
    hipStream_t s = rx_stream();
    const size_t np = cell.uNumParticles;
    if (!dGridD) {
        rx_set_error(2, "%s: null grid array", fn);
        return;
    }
    if (!np) {  // no particles: every lattice point has density 0 (the cell arrays may be unset)
        rx_fill<float>(dGridD, 0.0f, total, s);
        finish();
        return;
    }
    // slot 6: per slot the position with mu2 (16 B) and G with det(G) (48 B); per cell the smallest mu2 (4 B)
    char* ws = (char*)rx_scratch(6, 64 * np + 4 * (size_t)cell.uNumCells + 256);
    if (!ws) return;
    float4* rec = (float4*)ws;
    float4* gs = rec + np;
    float* cmu = (float*)(gs + 3 * np);
    k_aniso_prep<<<rx_blocks(np, kTpb), kTpb, 0, s>>>(dG, cell, rec, gs);
    k_aniso_cell_mu<<<rx_blocks(cell.uNumCells, kTpb), kTpb, 0, s>>>(rec, cell, cmu);
    k_aniso_grid_density<<<rx_blocks(total, kTpb), kTpb, 0, s>>>(dGridD, rec, gs, cmu, cell, aniso_consts(), make_uint3(nx, ny, nz),
                                                                 make_float3(x0, y0, z0), make_float3(dx, dy, dz));
    finish();
}

}  // extern "C"

// rx_sph.hip -- WCSPH and PBF neighbour kernels behind the extern "C" ABI (rx_cu_funcs.cuh:70-128).
//
// One thread per particle in sorted (cell) order.  The 27-cell neighbourhood is walked as 9 rows of
// consecutive cells: cells (x-1..x+1, y, z) have consecutive hashes, so their particles occupy one contiguous
// slot range of the sorted arrays, and lanes of a wavefront (consecutive sorted particles, hence neighbouring
// cells) stream overlapping ranges out of L1/L2.  Results are written in original particle order, which is
// the contract of these entry points.
#include <stdio.h>

#include "rx_common.h"
#include "../../include/rx_cu_funcs.h"

namespace {

constexpr int kTpb = 256;
// Library scratch slots (rx_scratch) of this file: the slot-order copies that the pair loops read.
constexpr int kSlotScalars = 13, kSlotBoundary = 14, kSlotVelDens = 15;

// Visits the slots of all particles in the cells overlapped by [p - h, p + h] (the cell range of
// rx_sph_kernel.cu:266-278), row by row and in slot order.  Candidates go in batches of four: `load(j)` for all
// four (the position, and the original index where the pass needs it), then `more(j, a)` for all four (the
// loads that only pairs within range need, issued under that test), then `body(j, a, b)` in slot order.  Every
// load of a stage is in flight before the first of them is used, instead of one candidate at a time.
struct Nothing {};
__device__ inline Nothing no_more(uint, const float4&) { return {}; }

template <class Load, class More, class Body>
__device__ inline void for_each_candidate(float3 p, const SphK& k, const uint* __restrict__ cs, const uint* __restrict__ ce,
                                          Load&& load, More&& more, Body&& body) {
    int3 g0 = grid_pos(f3(p.x - k.h, p.y - k.h, p.z - k.h), k);
    int3 g1 = grid_pos(f3(p.x + k.h, p.y + k.h, p.z + k.h), k);
    int nx = g1.x - g0.x + 1;
    for (int z = g0.z; z <= g1.z; ++z)
        for (int y = g0.y; y <= g1.y; ++y) {
            uint j, e;
            row_range(cs, ce, grid_hash(make_int3(g0.x, y, z), k), nx, j, e);
            for (; j + 3 < e; j += 4) {
                auto a0 = load(j);
                auto a1 = load(j + 1);
                auto a2 = load(j + 2);
                auto a3 = load(j + 3);
                auto b0 = more(j, a0);
                auto b1 = more(j + 1, a1);
                auto b2 = more(j + 2, a2);
                auto b3 = more(j + 3, a3);
                body(j, a0, b0);
                body(j + 1, a1, b1);
                body(j + 2, a2, b2);
                body(j + 3, a3, b3);
            }
            for (; j < e; ++j) {
                auto a = load(j);
                body(j, a, more(j, a));
            }
        }
}

struct PosIdx {  // sorted position of a candidate and its original index
    float4 x;
    uint o;
};

__device__ inline float pow_n(float x, float n) { return (n == 4.0f) ? (x * x) * (x * x) : powf(x, n); }

// ------------------------------------------------------------------------------------------ WCSPH
// rx_sph_kernel.cu:216-288 (sphCalDensity)
__global__ __launch_bounds__(kTpb) void k_sph_density(float* __restrict__ dens, float* __restrict__ pres, rxParticleCell c, SphK k, int write_pres) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    float d = 0.0f;
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; }, no_more,
        [&](uint, const float4& x, Nothing) {
            float3 r = p - f3(x);
            float r2 = dot3(r, r);
            if (r2 <= k.h2) {
                float q = k.h2 - r2;
                d += k.mass * k.wpoly6 * q * q * q;
            }
        });
    uint o = c.dSortedIndex[i];
    dens[o] = d;
    if (write_pres) pres[o] = k.gas_k * (d - k.rho0);
}

// rx_sph_kernel.cu:394-488 (sphCalForces) with the CPU path's pair rule (rx_sph_cpu.cpp:584: j != i, r > 0).
// sdp: (rho, p) in slot order (k_slot_pairs).
__global__ __launch_bounds__(kTpb) void k_sph_force(const float2* __restrict__ sdp, float4* __restrict__ frc, rxParticleCell c, SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p0 = f3(c.dSortedPos[i]);
    float3 v0 = f3(c.dSortedVel[i]);
    uint o = c.dSortedIndex[i];
    const float2 own = sdp[i];
    float d0 = own.x;
    float prsi = own.y / (d0 * d0);
    float3 fp = f3(0, 0, 0), fv = f3(0, 0, 0);
    struct Pair {
        bool in;
        float4 v;
        float d, pr;
    };
    for_each_candidate(
        p0, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair b{};
            float3 rij = p0 - f3(x);
            float r2 = dot3(rij, rij);
            b.in = j != i && r2 <= k.h2 && r2 > 0.0f;
            if (b.in) {
                b.v = c.dSortedVel[j];
                const float2 t = sdp[j];
                b.d = t.x;
                b.pr = t.y;
            }
            return b;
        },
        [&](uint, const float4& x, const Pair& b) {
            if (!b.in) return;
            float3 rij = p0 - f3(x);
            float r2 = dot3(rij, rij);
            float ri = __builtin_amdgcn_rsqf(r2);
            float id = __builtin_amdgcn_rcpf(b.d);  // 1 / rho_j, rho_j > 0 (the particle's own term)
            float prsj = b.pr * id * id;
            float q = k.h - r2 * ri;
            fp += (k.mass * (prsi + prsj) * k.gwspiky * q * q * ri) * rij;
            fv += (k.mass * id * k.lwvisc * q) * (f3(b.v) - v0);
        });
    float3 a = (-1.0f) * fp + k.visc * fv + k.gravity;
    frc[o] = make_float4(a.x, a.y, a.z, 0.0f);
}

// rx_sph_kernel.cu:301-376 (sphCalNormal)
// sd: rho in slot order (k_slot_scalars).
__global__ __launch_bounds__(kTpb) void k_sph_normal(float4* __restrict__ nrm, const float* __restrict__ sd, rxParticleCell c, SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p0 = f3(c.dSortedPos[i]);
    float3 n = f3(0, 0, 0);
    struct Pair {
        bool in;
        float d;
    };
    for_each_candidate(
        p0, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair b{};
            float3 rij = p0 - f3(x);
            float r = sqrtf(dot3(rij, rij));
            b.in = j != i && r <= k.h && r > 0.0001f;
            if (b.in) b.d = sd[j];
            return b;
        },
        [&](uint, const float4& x, const Pair& b) {
            if (!b.in) return;
            float3 rij = p0 - f3(x);
            float q = k.h2 - dot3(rij, rij);
            n += (k.mass * __builtin_amdgcn_rcpf(b.d) * k.gwpoly6 * q * q) * rij;
        });
    float l = sqrtf(dot3(n, n));
    if (l > 0.0f) n = (1.0f / l) * n;
    nrm[c.dSortedIndex[i]] = make_float4(n.x, n.y, n.z, 0.0f);
}

// Wall rule of the CPU path (rxSolidBox::GetDistanceR, rx_sph_solid.cpp:244-262): keep the centre at least
// ParticleRadius inside the box; velocity is not modified.
__device__ inline float3 wall_clamp(float3 x, const SphK& k) {
    x.x = fminf(fmaxf(x.x, k.wall_lo.x), k.wall_hi.x);
    x.y = fminf(fmaxf(x.y, k.wall_lo.y), k.wall_hi.y);
    x.z = fminf(fmaxf(x.z, k.wall_lo.z), k.wall_hi.z);
    return x;
}

// rx_sph_kernel.cu:814-837 (sphIntegrate) with the CPU wall rule (rx_sph_cpu.cpp:1040-1077).
__global__ void k_sph_integrate(float4* __restrict__ pos, float4* __restrict__ vel, const float4* __restrict__ acc,
                                const int* __restrict__ attr, float dt, uint n, SphK k) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (attr && attr[i] == -1) return;
    float3 x = f3(pos[i]), v = f3(vel[i]), a = f3(acc[i]);
    v += dt * a;
    x += dt * v;
    x = wall_clamp(x, k);
    pos[i] = make_float4(x.x, x.y, x.z, pos[i].w);
    vel[i] = make_float4(v.x, v.y, v.z, vel[i].w);
}

// rx_sph_kernel.cu:904-940 (checkDelete / checkDeleteX)
__global__ void k_check_delete(float4* pos, float4* vel, int* attr, float3 mn, float3 mx, float xmax, int mode, float3 far, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (attr[i] == -1) return;
    float3 x = f3(pos[i]);
    bool del = mode == 0 ? ((x.x > mn.x && x.x < mx.x) && (x.y > mn.y && x.y < mx.y) && (x.z > mn.z && x.z < mx.z)) : (x.x > xmax);
    if (del) {
        pos[i] = make_float4(far.x, far.y, far.z, 0.0f);
        vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        attr[i] = -1;
    }
}

// rx_sph_kernel.cu:1134-1208 (detectNeighborParticles).  As in the reference the test radius is the kernel
// support h (params.EffectiveRadius); the radius argument of the entry point is not used.  The distance is
// computed without contraction so that membership at r == h is the plain IEEE result.
__global__ __launch_bounds__(kTpb) void k_detect_neighbors(int* __restrict__ nb, rxParticleCell c, SphK k, int nmax) {
#pragma clang fp contract(off)
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    uint o = c.dSortedIndex[i];
    float3 p = f3(c.dSortedPos[i]);
    int cnt = 0;
    int3 g0 = grid_pos(f3(p.x - k.h, p.y - k.h, p.z - k.h), k);
    int3 g1 = grid_pos(f3(p.x + k.h, p.y + k.h, p.z + k.h), k);
    bool full = false;
    for (int z = g0.z; z <= g1.z && !full; ++z)
        for (int y = g0.y; y <= g1.y && !full; ++y)
            for (int x = g0.x; x <= g1.x && !full; ++x) {
                uint gh = grid_hash(make_int3(x, y, z), k);
                uint s = c.dCellStart[gh];
                if (s == kEmpty) continue;
                uint e = c.dCellEnd[gh];
                for (uint j = s; j < e; ++j) {
                    float3 q = f3(c.dSortedPos[j]);
                    float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                    float r = sqrtf(dx * dx + dy * dy + dz * dz);
                    if (r <= k.h) {
                        nb[(size_t)o * nmax + cnt] = (int)c.dSortedIndex[j];
                        cnt++;
                        if (cnt > nmax - 1) break;
                    }
                }
                if (cnt > 29) full = true;  // 30 = maxNeight (rx_sph_kernel.cu:1184)
            }
    for (int t = cnt; t < nmax - 1; ++t) nb[(size_t)o * nmax + t] = -1;
    nb[(size_t)o * nmax + nmax - 1] = cnt;
}

// rx_sph_kernel.cu:1214-1290 (CalDistToNormalizedMassCenter + detectSurfaceParticles), kept as written
// (including its loop bound of count - 1 neighbours).
__global__ void k_detect_surface(const int* __restrict__ nb, int* __restrict__ surf, rxParticleCell c, const float4* __restrict__ pos,
                                 float radius, int nmax, SphK k) {
#pragma clang fp contract(off)
    uint idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= c.uNumParticles) return;
    uint o = c.dSortedIndex[idx];
    float3 p0 = f3(pos[o]);
    int nn = nb[(size_t)(o + 1) * nmax - 1];
    float3 sp = f3(0, 0, 0);
    float sm = 0.0f;
    for (int t = 0; t < nn - 1; ++t) {
        int j = nb[(size_t)o * nmax + t];
        float3 p1 = f3(pos[j]);
        sp += k.mass * (p0 - p1);
        sm += k.mass;
    }
    float d = 1.17549435e-38f;  // FLT_MIN
    if (sm > 0.0f) {
        sp = make_float3(sp.x / sm, sp.y / sm, sp.z / sm);
        d = sqrtf(sp.x * sp.x + sp.y * sp.y + sp.z * sp.z);
    }
    if (nn <= 3) {
        surf[o] = nn;
    } else if (surf[o] > 0) {
        surf[o] = (d < 0.25f * radius) ? 0 : nn;
    } else {
        surf[o] = (d > 0.35f * radius || d < 0.0f) ? nn : 0;
    }
}

// ------------------------------------------------------------------------------------------ PBF
// rx_pbdsph_kernel.cu:71-101
__global__ __launch_bounds__(kTpb) void k_pbf_ext_force(const float* __restrict__ dens, float4* __restrict__ frc, rxParticleCell c, SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p0 = f3(c.dSortedPos[i]);
    float3 v0 = f3(c.dSortedVel[i]);
    uint o = c.dSortedIndex[i];
    float d0 = dens[o];
    float3 fv = f3(0, 0, 0);
    struct Pair {
        bool in;
        float4 v;
    };
    for_each_candidate(
        p0, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair b{};
            float3 rij = p0 - f3(x);
            float r2 = dot3(rij, rij);
            b.in = j != i && r2 <= k.h2 && r2 > 0.0f;
            if (b.in) b.v = c.dSortedVel[j];
            return b;
        },
        [&](uint, const float4& x, const Pair& b) {
            if (!b.in) return;
            float3 rij = p0 - f3(x);
            float r2 = dot3(rij, rij);
            float r = r2 * __builtin_amdgcn_rsqf(r2);  // r2 > 0 (the pair test above)
            fv += (k.lwvisc * (k.h - r)) * (f3(b.v) - v0);
        });
    // rx_sph_pbd.cpp:676: viscosity uses the particle's own density
    float3 a = (k.visc * k.mass / d0) * fv + k.gravity;
    frc[o] = make_float4(a.x, a.y, a.z, 0.0f);
}

// rx_pbdsph_kernel.cu:253-303 with the self term of rx_sph_pbd.cpp:767-784:
// sd = sum_{k != i} |grad W_ik / rho0|^2 + |sum_j grad W_ij / rho0|^2.
__global__ __launch_bounds__(kTpb) void k_pbf_lambda(float* __restrict__ dens, float* __restrict__ scl, float eps, rxParticleCell c, SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    float d = 0.0f, sd = 0.0f;
    float3 gs = f3(0, 0, 0);
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; }, no_more,
        [&](uint, const float4& x, Nothing) {
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            if (r2 <= k.h2) {
                float q2 = k.h2 - r2;
                d += k.mass * k.wpoly6 * q2 * q2 * q2;
                if (r2 > 0.0f) {
                    float ri = __builtin_amdgcn_rsqf(r2);
                    float q = k.h - r2 * ri;
                    float3 g = (k.gwspiky * k.inv_rho0 * q * q * ri) * rij;
                    gs += g;
                    sd += dot3(g, g);
                }
            }
        });
    sd += dot3(gs, gs);
    float C = d * k.inv_rho0 - 1.0f;
    uint o = c.dSortedIndex[i];
    scl[o] = -C / (sd + eps);
    dens[o] = d;
}

// Per-particle values in slot (sorted) order, so that a pair loop reads its neighbours' values from contiguous
// slots instead of gathering them by original index: one gather per particle instead of one per pair.
__global__ void k_slot_scalars(const float* __restrict__ src, const uint* __restrict__ idx, float* __restrict__ dst, uint n) {
    uint i = xcd_thread();
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_slot_pairs(const float* __restrict__ a, const float* __restrict__ b, const uint* __restrict__ idx,
                             float2* __restrict__ dst, uint n) {
    uint i = xcd_thread();
    if (i < n) {
        const uint o = idx[i];
        dst[i] = make_float2(a[o], b[o]);
    }
}
__global__ void k_slot_vel_dens(const float4* __restrict__ vel, const float* __restrict__ dens, const uint* __restrict__ idx,
                                float4* __restrict__ dst, uint n) {
    uint i = xcd_thread();
    if (i < n) {
        const uint o = idx[i];
        const float4 v = vel[o];
        dst[i] = make_float4(v.x, v.y, v.z, dens[o]);
    }
}

// rx_pbdsph_kernel.cu:374-407 / rx_sph_pbd.cpp:925-1026.  ss: the multipliers in slot order (k_slot_scalars).
__global__ __launch_bounds__(kTpb) void k_pbf_dp(const float* __restrict__ ss, float4* __restrict__ dp, rxParticleCell c, SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    float si = ss[i];
    float dt2 = k.dt * k.dt;
    float3 acc = f3(0, 0, 0);
    struct Pair {
        bool in;
        float s;
    };
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair b{};
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            b.in = r2 <= k.h2 && r2 > 0.0f;
            if (b.in) b.s = ss[j];
            return b;
        },
        [&](uint, const float4& x, const Pair& b) {
            if (!b.in) return;
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            float ri = __builtin_amdgcn_rsqf(r2);
            float sc = 0.0f;
            if (k.ap) {
                float q2 = k.h2 - r2;
                float ww = k.wpoly6 * q2 * q2 * q2 * k.inv_ap_wq;
                sc = -k.ap_k * pow_n(ww, k.ap_n) * dt2;
            }
            float q = k.h - r2 * ri;
            acc += ((si + b.s + sc) * k.gwspiky * k.inv_rho0 * q * q * ri) * rij;
        });
    dp[c.dSortedIndex[i]] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

__global__ void k_add4(float4* __restrict__ a, const float4* __restrict__ b, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = a[i], d = b[i];
    a[i] = make_float4(x.x + d.x, x.y + d.y, x.z + d.z, x.w + d.w);
}

// rx_pbdsph_kernel.cu:1066-1078 (pbdsphUpdateVelocity): new_vel = (pos - new_pos)/dt.
__global__ void k_pbf_update_velocity(const float4* __restrict__ pos, const float4* __restrict__ new_pos, float4* __restrict__ new_vel, float dt, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float3 v = (1.0f / dt) * (f3(pos[i]) - f3(new_pos[i]));
    new_vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
}
// pbdsphUpdatePosition: new_vel = (new_pos - pos)/dt; pos = new_pos.
__global__ void k_pbf_update_position(float4* __restrict__ pos, const float4* __restrict__ new_pos, float4* __restrict__ new_vel, float dt, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float3 x1 = f3(new_pos[i]);
    float3 v = (1.0f / dt) * (x1 - f3(pos[i]));
    new_vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
    pos[i] = make_float4(x1.x, x1.y, x1.z, pos[i].w);
}

// rx_pbdsph_kernel.cu:1130-1162 (xsphVisocosity)
// sv: (v, rho) in slot order (k_slot_vel_dens).
__global__ __launch_bounds__(kTpb) void k_xsph(const float4* __restrict__ sv, float4* __restrict__ nvel, float cx, rxParticleCell c,
                                               SphK k) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    uint o = c.dSortedIndex[i];
    float3 v0 = f3(sv[i]);
    float3 acc = f3(0, 0, 0);
    struct Pair {
        bool in;
        float4 v;
    };
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair b{};
            float3 rij = p - f3(x);
            b.in = dot3(rij, rij) <= k.h2;
            if (b.in) b.v = sv[j];
            return b;
        },
        [&](uint, const float4& x, const Pair& b) {
            if (!b.in) return;
            float3 rij = p - f3(x);
            float q = k.h2 - dot3(rij, rij);
            acc += (k.mass * __builtin_amdgcn_rcpf(b.v.w) * k.wpoly6 * q * q * q) * (f3(b.v) - v0);
        });
    float3 v = v0 + cx * acc;
    nvel[o] = make_float4(v.x, v.y, v.z, 0.0f);
}

// rx_pbdsph_kernel.cu:948-971 (pbdsphIntegrate) with the CPU wall rule.
__global__ void k_pbf_integrate(const float4* __restrict__ pos, const float4* __restrict__ vel, const float4* __restrict__ acc,
                                const int* __restrict__ attr, float4* __restrict__ npos, float4* __restrict__ nvel, float dt, uint n, SphK k) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x4 = pos[i];
    if (attr && attr[i] == -1) {
        npos[i] = x4;
        nvel[i] = vel[i];
        return;
    }
    float3 x = f3(x4), v = f3(vel[i]), a = f3(acc[i]);
    v += dt * a;
    x += dt * v;
    x = wall_clamp(x, k);
    npos[i] = make_float4(x.x, x.y, x.z, x4.w);
    nvel[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// rx_pbdsph_kernel.cu:431-441
__global__ void k_dens_fluct(float* __restrict__ err, const float* __restrict__ dens, float r0, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float e = dens[i] - r0;
    err[i] = (e >= 0.0f ? e : 0.0f) / r0;
}

// Max positive and mean absolute relative density error: per-block partials, then one block.
__global__ __launch_bounds__(256) void k_err_partial(const float* __restrict__ dens, float r0, uint n, float* __restrict__ pmax, double* __restrict__ psum) {
    __shared__ float smax[256];
    __shared__ double ssum[256];
    float mx = 0.0f;
    double sm = 0.0;
    for (uint i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float e = (dens[i] - r0) / r0;
        mx = fmaxf(mx, e);
        sm += fabs((double)e);
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = smax[0];
        psum[blockIdx.x] = ssum[0];
    }
}
__global__ __launch_bounds__(256) void k_err_final(const float* __restrict__ pmax, const double* __restrict__ psum, int nb, uint n, float* __restrict__ out) {
    __shared__ float smax[256];
    __shared__ double ssum[256];
    float mx = 0.0f;
    double sm = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        mx = fmaxf(mx, pmax[i]);
        sm += psum[i];
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smax[0];
        out[1] = (float)(n ? ssum[0] / n : 0.0);
    }
}



// ------------------------------------------------------------------------------------------ polygon obstacles
__device__ inline float3 cross3(float3 a, float3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// Segment P0 -> P1 against triangle (V0, V1, V2), entered from the side its normal (V1-V0) x (V2-V0) faces
// (intersectSegmentTriangle, rx_cu_common.cu:891-960, return value 1).  I = the crossing point on the plane,
// n = the unit normal.  Degenerate triangles and segments parallel to the plane do not hit.
__device__ inline bool segment_hits_triangle(float3 P0, float3 P1, float3 V0, float3 V1, float3 V2, float3& I, float3& n) {
    float3 u = V1 - V0, v = V2 - V0;
    float3 c = cross3(u, v);
    float cl = sqrtf(dot3(c, c));
    if (!(cl > 0.0f)) return false;
    n = (1.0f / cl) * c;
    float3 dir = P1 - P0;
    float a = dot3(n, P0 - V0), b = dot3(n, dir);
    if (fabsf(b) < 1e-10f) return false;
    float r = -a / b;
    if (r < 0.0f || fabsf(a) > fabsf(b) || b > 0.0f) return false;
    I = P0 + r * dir;
    float uu = dot3(u, u), uv = dot3(u, v), vv = dot3(v, v);
    float3 w = I - V0;
    float wu = dot3(w, u), wv = dot3(w, v);
    float D = uv * uv - uu * vv;
    float s = (uv * wv - vv * wu) / D;
    if (s < 0.0f || s > 1.0f) return false;
    float t = (uv * wu - uu * wv) / D;
    if (t < 0.0f || s + t > 1.0f) return false;
    return true;
}

// rxSPH::calCollisionPolygon (rx_sph_cpu.cpp:956-993) over the triangles listed in one cell of the polygon grid
// (rxParticleCell::dPolyCellStart/End/dSortedPolyIdx): on a hit the new position is put back at the crossing
// point plus the reflected velocity times dt d / l, and the velocity gains the reflection
// vr = -(1 + e') n (n.v), e' = e d / (dt |v|) (d = penetration depth, l = step length, e = Restitution).
__device__ inline void collide_cell_polygons(uint hsh, float3 x0, float3& x1, float3& v, const float3* __restrict__ vrts,
                                             const int3* __restrict__ tris, int ntri, const rxParticleCell& c, float dt, float e) {
    uint s = c.dPolyCellStart[hsh];
    if (s == kEmpty) return;
    uint end = c.dPolyCellEnd[hsh];
    for (uint j = s; j < end; ++j) {
        uint t = c.dSortedPolyIdx[j];
        if (t >= (uint)ntri) continue;
        int3 tri = tris[t];
        float3 cp, n;
        if (!segment_hits_triangle(x0, x1, vrts[tri.x], vrts[tri.y], vrts[tri.z], cp, n)) continue;
        float3 pen = x1 - cp;
        float d = sqrtf(dot3(pen, pen));
        float res = (e > 0.0f) ? e * fabsf(d) / (dt * sqrtf(dot3(v, v))) : 0.0f;
        float3 vr = (-(1.0f + res) * dot3(n, v)) * n;
        float3 step = x1 - x0;
        float l = sqrtf(dot3(step, step));
        x1 = cp + (dt * d / l) * vr;
        v += vr;
    }
}

// rxSPH::integrate with polygons (rx_sph_cpu.cpp:1040-1077; GPU forms sphIntegrateWithPolygon,
// rx_sph_kernel.cu:851-895, and pbdsphIntegrateWithPolygon, rx_pbdsph_kernel.cu:985-1030): symplectic Euler, the
// triangles of the old position's cell, then those of the new position's cell when it differs, then the walls.
// Results go to (npos, nvel); attr == -1 particles are left as they are (copied when npos != pos).
__global__ void k_integrate_polygon(const float4* pos, const float4* vel, const float4* __restrict__ acc, const int* __restrict__ attr,
                                    float4* npos, float4* nvel, const float3* __restrict__ vrts, const int3* __restrict__ tris,
                                    int ntri, float dt, rxParticleCell c, SphK k, float e, int keep_vel_w) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.uNumParticles) return;
    float4 x4 = pos[i], v4 = vel[i];
    if (attr && attr[i] == -1) {
        npos[i] = x4;
        nvel[i] = v4;
        return;
    }
    float3 x0 = f3(x4), v = f3(v4) + dt * f3(acc[i]);
    float3 x = x0 + dt * v;
    if (ntri > 0 && c.dPolyCellStart && c.dPolyCellEnd && c.dSortedPolyIdx) {
        uint h0 = grid_hash(grid_pos(x0, k), k);
        collide_cell_polygons(h0, x0, x, v, vrts, tris, ntri, c, dt, e);
        uint h1 = grid_hash(grid_pos(x, k), k);
        if (h1 != h0) collide_cell_polygons(h1, x0, x, v, vrts, tris, ntri, c, dt, e);
    }
    x = wall_clamp(x, k);
    npos[i] = make_float4(x.x, x.y, x.z, x4.w);
    nvel[i] = make_float4(v.x, v.y, v.z, keep_vel_w ? v4.w : 0.0f);
}

// ------------------------------------------------------------------------------------------ density on a grid
// sphCalDensityInGrid / pbdsphCalDensityInGrid / sphCalDensityInGridIceMesh (rx_sph_kernel.cu:1039-1126,
// rx_pbdsph_kernel.cu:1263-1300): rho at the nodes gmin + (x, y, z) glen of an nx x ny x nz lattice, stored at
// x + y nx + z nx ny.  With `ice`, particles whose flag is negative are left out.
__global__ __launch_bounds__(kTpb) void k_grid_density(float* __restrict__ out, rxParticleCell c, SphK k, uint3 gnum, float3 gmin,
                                                       float3 glen, const float* __restrict__ ice) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)gnum.x * gnum.y * gnum.z;
    if (idx >= total) return;
    uint gx = (uint)(idx % gnum.x), gy = (uint)((idx / gnum.x) % gnum.y), gz = (uint)(idx / ((size_t)gnum.x * gnum.y));
    float3 p = f3(gmin.x + gx * glen.x, gmin.y + gy * glen.y, gmin.z + gz * glen.z);
    struct Pair {
        bool in;
    };
    float d = 0.0f;
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return PosIdx{c.dSortedPos[j], c.dSortedIndex[j]}; },
        [&](uint, const PosIdx& a) {
            Pair q{true};
            if (ice) q.in = !(ice[a.o] < 0.0f);
            return q;
        },
        [&](uint, const PosIdx& a, const Pair& q) {
            float3 r = p - f3(a.x);
            float r2 = dot3(r, r);
            if (q.in && r2 <= k.h2) {
                float w = k.h2 - r2;
                d += k.mass * k.wpoly6 * w * w * w;
            }
        });
    out[idx] = d;
}

// ------------------------------------------------------------------------------------------ boundary particles
// Boundary particles (Akinci et al. 2012, "Versatile rigid-fluid coupling for incompressible SPH") sit in a
// second sorted grid whose geometry is rxSimParams::WorldOriginB / CellWidthB / GridSizeB
// (calcGridPosB/calcGridHashB, rx_cu_common.cu:155-178).  Each carries a volume V_b; it contributes
// rho0 V_b W to a fluid density and acts on the fluid as a particle of mass rho0 V_b.
SphK bgrid_consts(const rxSimParams& p) {
    SphK k = rx_make_consts(p);
    k.origin = p.WorldOriginB;
    k.cellw = p.CellWidthB;
    k.gx = (int)p.GridSizeB.x;
    k.gy = (int)p.GridSizeB.y;
    k.gz = (int)p.GridSizeB.z;
    k.ncells = p.NumCellsB;
    return k;
}

// rx_sph_kernel.cu:500-564 (sphCalBoundaryVolume): V_b = m / sum_{b'} m W(r_bb') over boundary particles,
// itself included (rx_sph_cpu.cpp:667-693).
__global__ __launch_bounds__(kTpb) void k_boundary_volume(float* __restrict__ vol, rxParticleCell b, SphK kb) {
    uint i = xcd_thread();
    if (i >= b.uNumParticles) return;
    float3 p = f3(b.dSortedPos[i]);
    float mw = 0.0f;
    for_each_candidate(
        p, kb, b.dCellStart, b.dCellEnd, [&](uint j) { return b.dSortedPos[j]; }, no_more,
        [&](uint, const float4& x, Nothing) {
            float3 r = p - f3(x);
            float r2 = dot3(r, r);
            if (r2 <= kb.h2) {
                float q = kb.h2 - r2;
                mw += kb.mass * kb.wpoly6 * q * q * q;
            }
        });
    vol[b.dSortedIndex[i]] = kb.mass / mw;
}

// Boundary term of a fluid density at p: sum_b rho0 V_b W(r).
__device__ inline float boundary_density(float3 p, const float* __restrict__ vs, const rxParticleCell& b, const SphK& kb) {
    struct Pair {
        bool in;
        float v;
    };
    float d = 0.0f;
    for_each_candidate(
        p, kb, b.dCellStart, b.dCellEnd, [&](uint j) { return b.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair q{};
            float3 r = p - f3(x);
            q.in = dot3(r, r) <= kb.h2;
            if (q.in) q.v = vs[j];
            return q;
        },
        [&](uint, const float4& x, const Pair& q) {
            if (!q.in) return;
            float3 r = p - f3(x);
            float w = kb.h2 - dot3(r, r);
            d += kb.rho0 * q.v * kb.wpoly6 * w * w * w;
        });
    return d;
}

// rx_sph_kernel.cu:575-650 (sphCalBoundaryDensity): rho_i += boundary term, p_i = k (rho_i - rho0); original order.
__global__ __launch_bounds__(kTpb) void k_sph_boundary_density(float* __restrict__ dens, float* __restrict__ pres,
                                                               const float4* __restrict__ pos, const float* __restrict__ vs,
                                                               rxParticleCell b, SphK kb, uint n, int write_pres) {
    uint i = xcd_thread();
    if (i >= n) return;
    float d = boundary_density(f3(pos[i]), vs, b, kb) + dens[i];
    dens[i] = d;
    if (write_pres) pres[i] = kb.gas_k * (d - kb.rho0);
}

// rx_sph_kernel.cu:662-727 (sphCalBoundaryForce): a_i += sum_b -rho0 V_b (p_i / rho_i^2) grad W_spiky(r_ib),
// for 1e-4 < r <= h.  This is the GPU form: an acceleration like the rest of dFrc.  The CPU loop
// (rx_sph_cpu.cpp:617-651) divides the same sum by the particle mass, which is not an acceleration.
__global__ __launch_bounds__(kTpb) void k_sph_boundary_force(const float* __restrict__ dens, const float* __restrict__ pres,
                                                             const float4* __restrict__ pos, const float* __restrict__ vs,
                                                             float4* __restrict__ frc, rxParticleCell b, SphK kb, uint n) {
    uint i = xcd_thread();
    if (i >= n) return;
    float3 p = f3(pos[i]);
    float d0 = dens[i];
    float prsi = pres[i] / (d0 * d0);
    struct Pair {
        bool in;
        float v;
    };
    float3 f = f3(0, 0, 0);
    for_each_candidate(
        p, kb, b.dCellStart, b.dCellEnd, [&](uint j) { return b.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair q{};
            float3 rij = p - f3(x);
            float r = sqrtf(dot3(rij, rij));
            q.in = r <= kb.h && r > 0.0001f;
            if (q.in) q.v = vs[j];
            return q;
        },
        [&](uint, const float4& x, const Pair& q) {
            if (!q.in) return;
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            float ri = __builtin_amdgcn_rsqf(r2);  // r > 1e-4 (the pair test above)
            float w = kb.h - r2 * ri;
            f += (-kb.rho0 * q.v * prsi * kb.gwspiky * w * w * ri) * rij;
        });
    float4 a = frc[i];
    frc[i] = make_float4(a.x + f.x, a.y + f.y, a.z + f.z, a.w);
}

// Boundary terms of a PBF multiplier at p: the density sum (as above) and sum_b |(rho0 V_b / m) grad W / rho0|^2
// (rx_pbdsph_kernel.cu:533-560, rx_sph_pbd.cpp:790-812).
__device__ inline void boundary_lambda_terms(float3 p, const float* __restrict__ vs, const rxParticleCell& b, const SphK& kb,
                                             float& d, float& sd) {
    const float gv = kb.rho0 / kb.mass * kb.gwspiky * kb.inv_rho0;  // gradient factor per unit boundary volume
    struct Pair {
        bool in;
        float v;
    };
    for_each_candidate(
        p, kb, b.dCellStart, b.dCellEnd, [&](uint j) { return b.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair q{};
            float3 r = p - f3(x);
            q.in = dot3(r, r) <= kb.h2;
            if (q.in) q.v = vs[j];
            return q;
        },
        [&](uint, const float4& x, const Pair& q) {
            if (!q.in) return;
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            float w = kb.h2 - r2;
            d += kb.rho0 * q.v * kb.wpoly6 * w * w * w;
            if (r2 > 0.0f) {
                float ri = __builtin_amdgcn_rsqf(r2);
                float t = kb.h - r2 * ri;
                float3 g = (q.v * gv * t * t * ri) * rij;
                sd += dot3(g, g);
            }
        });
}

// rx_pbdsph_kernel.cu:493-520 (pbdsphCalBoundaryDensity): rho_i += boundary term; original order.
// (k_sph_boundary_density with write_pres = 0.)

// rx_pbdsph_kernel.cu:577-657 (pbdsphCalScalingFactorWithBoundary) with the fluid terms of k_pbf_lambda (self
// term included, rx_sph_pbd.cpp:699-822): rho_i and sd_i gain the boundary terms.
__global__ __launch_bounds__(kTpb) void k_pbf_lambda_wb(float* __restrict__ dens, float* __restrict__ scl, float eps, rxParticleCell c,
                                                        SphK k, const float* __restrict__ vs, rxParticleCell b, SphK kb) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    float d = 0.0f, sd = 0.0f;
    float3 gs = f3(0, 0, 0);
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; }, no_more,
        [&](uint, const float4& x, Nothing) {
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            if (r2 <= k.h2) {
                float q2 = k.h2 - r2;
                d += k.mass * k.wpoly6 * q2 * q2 * q2;
                if (r2 > 0.0f) {
                    float ri = __builtin_amdgcn_rsqf(r2);
                    float q = k.h - r2 * ri;
                    float3 g = (k.gwspiky * k.inv_rho0 * q * q * ri) * rij;
                    gs += g;
                    sd += dot3(g, g);
                }
            }
        });
    sd += dot3(gs, gs);
    boundary_lambda_terms(p, vs, b, kb, d, sd);
    float C = d * k.inv_rho0 - 1.0f;
    uint o = c.dSortedIndex[i];
    scl[o] = -C / (sd + eps);
    dens[o] = d;
}

// rx_pbdsph_kernel.cu:666-745 (pbdsphCalBoundaryScalingFactor) with the CPU rule of rx_sph_pbd.cpp:825-906:
// lambda_b = -C_b / (sd_b + eps), C_b = (sum_fluid m W + sum_b' rho0 V_b' W) / rho0 - 1,
// sd_b = sum_fluid |grad W / rho0|^2 + sum_b' |(rho0 V_b' / m) grad W / rho0|^2 (r > 0).  Every fluid particle
// within h counts (the GPU kernel skips the fluid slot whose number equals the boundary slot's).
__global__ __launch_bounds__(kTpb) void k_pbf_boundary_lambda(float* __restrict__ sclb, float eps, rxParticleCell c, SphK k,
                                                              const float* __restrict__ vs, rxParticleCell b, SphK kb) {
    uint i = xcd_thread();
    if (i >= b.uNumParticles) return;
    float3 p = f3(b.dSortedPos[i]);
    float d = 0.0f, sd = 0.0f;
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; }, no_more,
        [&](uint, const float4& x, Nothing) {
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            if (r2 <= k.h2) {
                float q2 = k.h2 - r2;
                d += k.mass * k.wpoly6 * q2 * q2 * q2;
                if (r2 > 0.0f) {
                    float ri = __builtin_amdgcn_rsqf(r2);
                    float q = k.h - r2 * ri;
                    float3 g = (k.gwspiky * k.inv_rho0 * q * q * ri) * rij;
                    sd += dot3(g, g);
                }
            }
        });
    boundary_lambda_terms(p, vs, b, kb, d, sd);
    float C = d * k.inv_rho0 - 1.0f;
    sclb[b.dSortedIndex[i]] = -C / (sd + eps);
}

// rx_pbdsph_kernel.cu:760-860 (pbdsphPositionCorrectionWithBoundary), rx_sph_pbd.cpp:978-1018: the fluid terms of
// k_pbf_dp plus sum_b (lambda_i + lambda_b + s_corr) grad W_spiky / rho0 with s_corr's W scaled by rho0 V_b / m.
// ss: the fluid multipliers in slot order; sb: (lambda_b, V_b) in the boundary grid's slot order.
__global__ __launch_bounds__(kTpb) void k_pbf_dp_wb(const float* __restrict__ ss, float4* __restrict__ dp, rxParticleCell c, SphK k,
                                                    const float2* __restrict__ sb, rxParticleCell b, SphK kb) {
    uint i = xcd_thread();
    if (i >= c.uNumParticles) return;
    float3 p = f3(c.dSortedPos[i]);
    float si = ss[i];
    float dt2 = k.dt * k.dt;
    const float wv = kb.rho0 / kb.mass * kb.wpoly6 * kb.inv_ap_wq;  // s_corr's W ratio per unit boundary volume
    float3 acc = f3(0, 0, 0);
    struct Pair {
        bool in;
        float s, v;
    };
    for_each_candidate(
        p, k, c.dCellStart, c.dCellEnd, [&](uint j) { return c.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair q{};
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            q.in = r2 <= k.h2 && r2 > 0.0f;
            if (q.in) q.s = ss[j];
            return q;
        },
        [&](uint, const float4& x, const Pair& q) {
            if (!q.in) return;
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            float ri = __builtin_amdgcn_rsqf(r2);
            float sc = 0.0f;
            if (k.ap) {
                float q2 = k.h2 - r2;
                float ww = k.wpoly6 * q2 * q2 * q2 * k.inv_ap_wq;
                sc = -k.ap_k * pow_n(ww, k.ap_n) * dt2;
            }
            float t = k.h - r2 * ri;
            acc += ((si + q.s + sc) * k.gwspiky * k.inv_rho0 * t * t * ri) * rij;
        });
    for_each_candidate(
        p, kb, b.dCellStart, b.dCellEnd, [&](uint j) { return b.dSortedPos[j]; },
        [&](uint j, const float4& x) {
            Pair q{};
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            q.in = r2 <= kb.h2 && r2 > 0.0f;
            if (q.in) {
                const float2 t = sb[j];
                q.s = t.x;
                q.v = t.y;
            }
            return q;
        },
        [&](uint, const float4& x, const Pair& q) {
            if (!q.in) return;
            float3 rij = p - f3(x);
            float r2 = dot3(rij, rij);
            float ri = __builtin_amdgcn_rsqf(r2);
            float sc = 0.0f;
            if (kb.ap) {
                float q2 = kb.h2 - r2;
                float ww = q.v * wv * q2 * q2 * q2;
                sc = -kb.ap_k * pow_n(ww, kb.ap_n) * dt2;
            }
            float t = kb.h - r2 * ri;
            acc += ((si + q.s + sc) * kb.gwspiky * kb.inv_rho0 * t * t * ri) * rij;
        });
    dp[c.dSortedIndex[i]] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

bool need_params(const char* fn) {
    if (!rx_params_set()) {
        rx_set_error(3, "%s: CuSetParameters must be called first", fn);
        return false;
    }
    return true;
}

// The boundary grid of `b` must match rxSimParams::GridSizeB (its cell arrays are indexed by that geometry).
bool boundary_grid_ok(const rxParticleCell& b, const char* fn) {
    const rxSimParams& p = rx_params();
    size_t cells = (size_t)p.GridSizeB.x * p.GridSizeB.y * p.GridSizeB.z;
    if (cells == 0 || cells != b.uNumCells || !b.dCellStart || !b.dCellEnd || (b.uNumParticles && (!b.dSortedPos || !b.dSortedIndex))) {
        rx_set_error(2, "%s: boundary grid (GridSizeB %ux%ux%u) does not match the cell data (%u cells)", fn, p.GridSizeB.x,
                     p.GridSizeB.y, p.GridSizeB.z, b.uNumCells);
        return false;
    }
    return true;
}

void finish(const char* fn) {
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(rx_stream()));
}

}  // namespace

// ============================================================================================ extern "C"
extern "C" {

void CuSphDensity(float* dDens, float* dPres, rxParticleCell cell) {
    if (!need_params("CuSphDensity") || !cell.uNumParticles) return;
    k_sph_density<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dDens, dPres, cell, rx_make_consts(rx_params()), 1);
    finish("CuSphDensity");
}

void CuSphForces(float* dDens, float* dPres, float* dFrc, rxParticleCell cell, float dt) {
    if (!need_params("CuSphForces") || !cell.uNumParticles) return;
    const uint n = cell.uNumParticles;
    float2* sdp = (float2*)rx_scratch(kSlotScalars, sizeof(float2) * n);
    if (!sdp) return;
    k_slot_pairs<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(dDens, dPres, cell.dSortedIndex, sdp, n);
    k_sph_force<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(sdp, (float4*)dFrc, cell, rx_make_consts(rx_params()));
    finish("CuSphForces");
}

void CuSphNormal(float* dNrms, float* dDens, rxParticleCell cell) {
    if (!need_params("CuSphNormal") || !cell.uNumParticles) return;
    const uint n = cell.uNumParticles;
    float* sd = (float*)rx_scratch(kSlotScalars, sizeof(float) * n);
    if (!sd) return;
    k_slot_scalars<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(dDens, cell.dSortedIndex, sd, n);
    k_sph_normal<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>((float4*)dNrms, sd, cell, rx_make_consts(rx_params()));
    finish("CuSphNormal");
}

void CuSphIntegrate(float* pos, float* vel, float* frc, float* dens, int* attr, float dt, uint numParticles) {
    if (!need_params("CuSphIntegrate") || !numParticles) return;
    k_sph_integrate<<<rx_blocks(numParticles, kTpb), kTpb, 0, rx_stream()>>>((float4*)pos, (float4*)vel, (const float4*)frc, attr, dt,
                                                                            numParticles, rx_make_consts(rx_params()));
    finish("CuSphIntegrate");
}

void CuSphCheckDelete(float* pos, float* vel, int* attr, float minp[3], float maxp[3], float far_point[3], uint nprts) {
    if (!nprts) return;
    k_check_delete<<<rx_blocks(nprts, kTpb), kTpb, 0, rx_stream()>>>((float4*)pos, (float4*)vel, attr, f3(minp[0], minp[1], minp[2]),
                                                                    f3(maxp[0], maxp[1], maxp[2]), 0.0f, 0,
                                                                    f3(far_point[0], far_point[1], far_point[2]), nprts);
    finish("CuSphCheckDelete");
}
void CuSphCheckDeleteX(float* pos, float* vel, int* attr, float xmax, float farpoint[3], uint nprts) {
    if (!nprts) return;
    k_check_delete<<<rx_blocks(nprts, kTpb), kTpb, 0, rx_stream()>>>((float4*)pos, (float4*)vel, attr, f3(0, 0, 0), f3(0, 0, 0), xmax, 1,
                                                                    f3(farpoint[0], farpoint[1], farpoint[2]), nprts);
    finish("CuSphCheckDeleteX");
}
void CuPbdSphCheckDelete(float* pos, float* vel, int* attr, float minp[3], float maxp[3], float far_point[3], uint nprts) {
    CuSphCheckDelete(pos, vel, attr, minp, maxp, far_point, nprts);
}
void CuPbdSphCheckDeleteX(float* pos, float* vel, int* attr, float xmax, float farpoint[3], uint nprts) {
    CuSphCheckDeleteX(pos, vel, attr, xmax, farpoint, nprts);
}

void CuSphDetectNeighborParticles(int* neights, rxParticleCell celldata, float radius, int prtNum) {
    if (!need_params("CuSphDetectNeighborParticles") || prtNum <= 0) return;
    celldata.uNumParticles = (uint)prtNum;
    k_detect_neighbors<<<rx_blocks(prtNum, kTpb), kTpb, 0, rx_stream()>>>(neights, celldata, rx_make_consts(rx_params()), 30);
    finish("CuSphDetectNeighborParticles");
}

void CuSphDetectSurfaceParticles(int* neights, int* surface, rxParticleCell celldata, float* pos, float radius) {
    if (!need_params("CuSphDetectSurfaceParticles") || !celldata.uNumParticles) return;
    k_detect_surface<<<rx_blocks(celldata.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(neights, surface, celldata, (const float4*)pos,
                                                                                       radius, 30, rx_make_consts(rx_params()));
    finish("CuSphDetectSurfaceParticles");
}

void CuPbdSphDensity(float* dDens, rxParticleCell cell) {
    if (!need_params("CuPbdSphDensity") || !cell.uNumParticles) return;
    k_sph_density<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dDens, nullptr, cell, rx_make_consts(rx_params()), 0);
    finish("CuPbdSphDensity");
}

void CuPbdSphExternalForces(float* dDens, float* dFrc, rxParticleCell cell, float dt) {
    if (!need_params("CuPbdSphExternalForces") || !cell.uNumParticles) return;
    k_pbf_ext_force<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dDens, (float4*)dFrc, cell, rx_make_consts(rx_params()));
    finish("CuPbdSphExternalForces");
}

void CuPbdSphScalingFactor(float* dPos, float* dDens, float* dScl, float eps, rxParticleCell cell) {
    if (!need_params("CuPbdSphScalingFactor") || !cell.uNumParticles) return;
    k_pbf_lambda<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dDens, dScl, eps, cell, rx_make_consts(rx_params()));
    finish("CuPbdSphScalingFactor");
}

void CuPbdSphPositionCorrection(float* dPos, float* dScl, float* dDp, rxParticleCell cell) {
    if (!need_params("CuPbdSphPositionCorrection") || !cell.uNumParticles) return;
    const uint n = cell.uNumParticles;
    float* ss = (float*)rx_scratch(kSlotScalars, sizeof(float) * n);
    if (!ss) return;
    k_slot_scalars<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(dScl, cell.dSortedIndex, ss, n);
    k_pbf_dp<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(ss, (float4*)dDp, cell, rx_make_consts(rx_params()));
    finish("CuPbdSphPositionCorrection");
}

void CuPbdSphCorrectPosition(float* dPos, float* dDp, uint nprts) {
    if (!nprts) return;
    k_add4<<<rx_blocks(nprts, kTpb), kTpb, 0, rx_stream()>>>((float4*)dPos, (const float4*)dDp, nprts);
    finish("CuPbdSphCorrectPosition");
}

void CuPbdSphUpdatePosition(float* pos, float* new_pos, float* new_vel, float dt, uint nprts) {
    if (!nprts) return;
    k_pbf_update_position<<<rx_blocks(nprts, kTpb), kTpb, 0, rx_stream()>>>((float4*)pos, (const float4*)new_pos, (float4*)new_vel, dt, nprts);
    finish("CuPbdSphUpdatePosition");
}

void CuPbdSphUpdateVelocity(float* pos, float* new_pos, float* new_vel, float dt, uint nprts) {
    if (!nprts) return;
    k_pbf_update_velocity<<<rx_blocks(nprts, kTpb), kTpb, 0, rx_stream()>>>((const float4*)pos, (const float4*)new_pos, (float4*)new_vel, dt, nprts);
    finish("CuPbdSphUpdateVelocity");
}

void CuXSphViscosity(float* dPos, float* dVel, float* dNewVel, float* dDens, float c, rxParticleCell cell) {
    if (!need_params("CuXSphViscosity") || !cell.uNumParticles) return;
    const uint n = cell.uNumParticles;
    float4* sv = (float4*)rx_scratch(kSlotVelDens, sizeof(float4) * n);
    if (!sv) return;
    k_slot_vel_dens<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>((const float4*)dVel, dDens, cell.dSortedIndex, sv, n);
    k_xsph<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(sv, (float4*)dNewVel, c, cell, rx_make_consts(rx_params()));
    finish("CuXSphViscosity");
}

float CuPbdSphCalDensityFluctuation(float* dErrScan, float* dErr, float* dDens, float rest_dens, uint nprts) {
    if (!nprts) return 0.0f;
    hipStream_t s = rx_stream();
    k_dens_fluct<<<rx_blocks(nprts, kTpb), kTpb, 0, s>>>(dErr, dDens, rest_dens, nprts);
    rx_scan_f32(dErr, dErrScan, nprts, s);
    float last[2] = {0, 0};
    RX_HIP(hipMemcpyAsync(&last[0], dErrScan + nprts - 1, sizeof(float), hipMemcpyDeviceToHost, s));
    RX_HIP(hipMemcpyAsync(&last[1], dErr + nprts - 1, sizeof(float), hipMemcpyDeviceToHost, s));
    RX_HIP(hipStreamSynchronize(s));
    return (last[0] + last[1]) / (float)nprts;
}

void CuPbdSphIntegrate(float* pos, float* vel, float* acc, int* attr, float* new_pos, float* new_vel, float dt, uint numParticles) {
    if (!need_params("CuPbdSphIntegrate") || !numParticles) return;
    k_pbf_integrate<<<rx_blocks(numParticles, kTpb), kTpb, 0, rx_stream()>>>((const float4*)pos, (const float4*)vel, (const float4*)acc, attr,
                                                                            (float4*)new_pos, (float4*)new_vel, dt, numParticles,
                                                                            rx_make_consts(rx_params()));
    finish("CuPbdSphIntegrate");
}

void CuPbdSphNormal(float* dNrms, float* dDens, rxParticleCell cell) { CuSphNormal(dNrms, dDens, cell); }

void CuPbdSphDensityErrorStats(float* dDens, float rest_dens, uint nprts, float* host_max_pos, float* host_mean_abs) {
    if (!nprts) {
        if (host_max_pos) *host_max_pos = 0.0f;
        if (host_mean_abs) *host_mean_abs = 0.0f;
        return;
    }
    hipStream_t s = rx_stream();
    int nb = (int)std::min<uint>(rx_blocks(nprts, 256), 1024);
    char* scr = (char*)rx_scratch(5, (sizeof(float) + sizeof(double)) * nb + 64);
    if (!scr) return;
    double* psum = (double*)scr;
    float* pmax = (float*)(scr + sizeof(double) * nb);
    float* out = pmax + nb;
    k_err_partial<<<nb, 256, 0, s>>>(dDens, rest_dens, nprts, pmax, psum);
    k_err_final<<<1, 256, 0, s>>>(pmax, psum, nb, nprts, out);
    float h[2];
    RX_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    RX_HIP(hipStreamSynchronize(s));
    if (host_max_pos) *host_max_pos = h[0];
    if (host_mean_abs) *host_mean_abs = h[1];
}


// ---------------------------------------------------------------- boundary particles (rx_cu_funcs.cuh:81-83, 131-135)
// rx_sph.cu:501-512.  As in the reference the volume uses rxSimParams::Mass; the mass argument is not read.
void CuSphBoundaryVolume(float* dVolB, float mass, rxParticleCell cell) {
    if (!need_params("CuSphBoundaryVolume") || !boundary_grid_ok(cell, "CuSphBoundaryVolume") || !cell.uNumParticles) return;
    k_boundary_volume<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dVolB, cell, bgrid_consts(rx_params()));
    finish("CuSphBoundaryVolume");
}

// dVolB in the boundary grid's slot order (library scratch), or null after a failed allocation.
float* boundary_volumes(const float* vol, const rxParticleCell& b) {
    const uint nb = b.uNumParticles;
    float* vs = (float*)rx_scratch(kSlotBoundary, sizeof(float) * (nb ? nb : 1));
    if (vs && nb) k_slot_scalars<<<rx_blocks(nb, kTpb), kTpb, 0, rx_stream()>>>(vol, b.dSortedIndex, vs, nb);
    return vs;
}

// rx_sph.cu:520-532
void CuSphBoundaryDensity(float* dDens, float* dPres, float* dPos, float* dVolB, rxParticleCell bcell, uint pnum) {
    if (!need_params("CuSphBoundaryDensity") || !boundary_grid_ok(bcell, "CuSphBoundaryDensity") || !pnum) return;
    const float* vs = boundary_volumes(dVolB, bcell);
    if (!vs) return;
    k_sph_boundary_density<<<rx_blocks(pnum, kTpb), kTpb, 0, rx_stream()>>>(dDens, dPres, (const float4*)dPos, vs, bcell,
                                                                            bgrid_consts(rx_params()), pnum, 1);
    finish("CuSphBoundaryDensity");
}

// rx_sph.cu:541-553
void CuSphBoundaryForces(float* dDens, float* dPres, float* dPos, float* dVolB, float* dFrc, rxParticleCell bcell, uint pnum) {
    if (!need_params("CuSphBoundaryForces") || !boundary_grid_ok(bcell, "CuSphBoundaryForces") || !pnum) return;
    const float* vs = boundary_volumes(dVolB, bcell);
    if (!vs) return;
    k_sph_boundary_force<<<rx_blocks(pnum, kTpb), kTpb, 0, rx_stream()>>>(dDens, dPres, (const float4*)dPos, vs, (float4*)dFrc,
                                                                          bcell, bgrid_consts(rx_params()), pnum);
    finish("CuSphBoundaryForces");
}

// rx_sph.cu:754-766
void CuPbdSphBoundaryDensity(float* dDens, float* dPos, float* dVolB, rxParticleCell bcell, uint pnum) {
    if (!need_params("CuPbdSphBoundaryDensity") || !boundary_grid_ok(bcell, "CuPbdSphBoundaryDensity") || !pnum) return;
    const float* vs = boundary_volumes(dVolB, bcell);
    if (!vs) return;
    k_sph_boundary_density<<<rx_blocks(pnum, kTpb), kTpb, 0, rx_stream()>>>(dDens, nullptr, (const float4*)dPos, vs, bcell,
                                                                            bgrid_consts(rx_params()), pnum, 0);
    finish("CuPbdSphBoundaryDensity");
}

// rx_sph.cu:779-815: the fluid multipliers and densities, then the boundary particles' multipliers (dSclB).
void CuPbdSphScalingFactorWithBoundary(float* dPos, float* dDens, float* dScl, float eps, rxParticleCell cell, float* dVolB,
                                       float* dSclB, rxParticleCell bcell) {
    if (!need_params("CuPbdSphScalingFactorWithBoundary") || !boundary_grid_ok(bcell, "CuPbdSphScalingFactorWithBoundary")) return;
    SphK k = rx_make_consts(rx_params()), kb = bgrid_consts(rx_params());
    const float* vs = boundary_volumes(dVolB, bcell);
    if (!vs) return;
    if (cell.uNumParticles)
        k_pbf_lambda_wb<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dDens, dScl, eps, cell, k, vs, bcell, kb);
    if (bcell.uNumParticles)
        k_pbf_boundary_lambda<<<rx_blocks(bcell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(dSclB, eps, cell, k, vs, bcell, kb);
    finish("CuPbdSphScalingFactorWithBoundary");
}

// rx_sph.cu:827-850
void CuPbdSphPositionCorrectionWithBoundary(float* dPos, float* dScl, float* dDp, rxParticleCell cell, float* dVolB, float* dSclB,
                                            rxParticleCell bcell) {
    if (!need_params("CuPbdSphPositionCorrectionWithBoundary") || !boundary_grid_ok(bcell, "CuPbdSphPositionCorrectionWithBoundary") ||
        !cell.uNumParticles)
        return;
    const uint n = cell.uNumParticles, nb = bcell.uNumParticles;
    float* ss = (float*)rx_scratch(kSlotScalars, sizeof(float) * n);
    float2* sb = (float2*)rx_scratch(kSlotBoundary, sizeof(float2) * (nb ? nb : 1));
    if (!ss || !sb) return;
    k_slot_scalars<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(dScl, cell.dSortedIndex, ss, n);
    if (nb) k_slot_pairs<<<rx_blocks(nb, kTpb), kTpb, 0, rx_stream()>>>(dSclB, dVolB, bcell.dSortedIndex, sb, nb);
    k_pbf_dp_wb<<<rx_blocks(n, kTpb), kTpb, 0, rx_stream()>>>(ss, (float4*)dDp, cell, rx_make_consts(rx_params()), sb, bcell,
                                                               bgrid_consts(rx_params()));
    finish("CuPbdSphPositionCorrectionWithBoundary");
}


// ---------------------------------------------------------------- polygon obstacles (rx_cu_funcs.cuh:74-75, 118-120)
// rx_sph.cu:307-320.  vrts: float3 per vertex, tris: 3 vertex indices per triangle, cell.dPolyCellStart/End/
// dSortedPolyIdx: the triangles listed per cell of the fluid grid (built by the caller).  In place.
void CuSphIntegrateWithPolygon(float* pos, float* vel, float* frc, float* dens, int* attr, float* vrts, int* tris, int tri_num,
                               float dt, rxParticleCell cell) {
    if (!need_params("CuSphIntegrateWithPolygon") || !cell.uNumParticles) return;
    k_integrate_polygon<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(
        (const float4*)pos, (const float4*)vel, (const float4*)frc, attr, (float4*)pos, (float4*)vel, (const float3*)vrts,
        (const int3*)tris, tri_num, dt, cell, rx_make_consts(rx_params()), rx_params().Restitution, 1);
    finish("CuSphIntegrateWithPolygon");
}

// rx_sph.cu:893-906.  Results in (new_pos, new_vel); pos/new_pos may alias.
void CuPbdSphIntegrateWithPolygon(float* pos, float* vel, float* acc, int* attr, float* new_pos, float* new_vel, float* vrts, int* tris,
                                  int tri_num, float dt, rxParticleCell cell) {
    if (!need_params("CuPbdSphIntegrateWithPolygon") || !cell.uNumParticles) return;
    k_integrate_polygon<<<rx_blocks(cell.uNumParticles, kTpb), kTpb, 0, rx_stream()>>>(
        (const float4*)pos, (const float4*)vel, (const float4*)acc, attr, (float4*)new_pos, (float4*)new_vel, (const float3*)vrts,
        (const int3*)tris, tri_num, dt, cell, rx_make_consts(rx_params()), rx_params().Restitution, 0);
    finish("CuPbdSphIntegrateWithPolygon");
}

// ---------------------------------------------------------------- density on a grid (rx_cu_funcs.cuh:86-91, 126)
static void grid_density(const char* fn, float* dGridD, const rxParticleCell& cell, int nx, int ny, int nz, float x0, float y0, float z0,
                         float dx, float dy, float dz, const float* ice) {
    if (!need_params(fn)) return;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    size_t total = (size_t)nx * ny * nz;
    k_grid_density<<<rx_blocks(total, kTpb), kTpb, 0, rx_stream()>>>(dGridD, cell, rx_make_consts(rx_params()),
                                                                     make_uint3((uint)nx, (uint)ny, (uint)nz), f3(x0, y0, z0),
                                                                     f3(dx, dy, dz), ice);
    finish(fn);
}

// rx_sph.cu:390-425
void CuSphGridDensity(float* dGridD, rxParticleCell cell, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                      float dz) {
    grid_density("CuSphGridDensity", dGridD, cell, nx, ny, nz, x0, y0, z0, dx, dy, dz, nullptr);
}
// rx_sph.cu:1054-1086
void CuPbdSphGridDensity(float* dGridD, rxParticleCell cell, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                         float dz) {
    grid_density("CuPbdSphGridDensity", dGridD, cell, nx, ny, nz, x0, y0, z0, dx, dy, dz, nullptr);
}
// rx_sph.cu:433-468: particles with bIceFlag < 0 (not ice) are left out.
void CuIceMeshMake(float* dGridD, rxParticleCell cell, int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy,
                   float dz, float* bIceFlag) {
    grid_density("CuIceMeshMake", dGridD, cell, nx, ny, nz, x0, y0, z0, dx, dy, dz, bIceFlag);
}

}  // extern "C"

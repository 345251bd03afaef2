// pbf_solver.hip -- fused position-based-fluids step (rxPBDSPH_GPU::Update, rx_sph_pbd_gpu.cpp:790-957).
//
// Data flow per step (state kept in cell order):
//   predict      density + viscosity at x_n (neighbour list of the previous step), x* = x + dt (v + dt a)
//   rebuild      stable counting-sort grid on x* (cell = h + skin), gather of the state
//   list+lambda  one stencil pass over the new grid: skin neighbour list (r <= h + skin) and lambda_0
//   K x          position correction (list) and lambda (list), ping-pong between two position buffers
//   velocity     v = (x* - x_n) / dt
//   XSPH         v += c sum (m/rho_j)(v_j - v_i) W (list)
// Every list pass reads the largest displacement since the list was built (kept on the device); while it is
// within skin/2 the list holds every pair within h.  Past that, the few particles that moved more than skin/2
// ("movers", listed by the correction pass) walk the build-time grid, and every other particle walks its list
// plus the nearby movers its list lacks; with too many movers every particle bounds displacements by the
// largest one.  The result never depends on the list being fresh.  No host synchronisation happens inside a
// step.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "fused_grid.h"
#include "pbf_solver.h"
#include "radix.h"
#include "mlp.h"
#include "../../include/rx_cu_funcs.h"
#include "../../include/rx_sph_host.h"

struct PbfK {
    SphK s;
    FGridK g;
    uint n;
    float rc2;        // (h + skin)^2, list cutoff
    float half_skin2; // (skin / 2)^2
    float skin;       // list cutoff minus h
    float wide2;      // largest displacement^2 the capped fallback walk still covers
    uint* fallback;   // counts the passes of particles that walked the grid instead of their list
    uint* bad;        // stats slots ST_BAD and ST_RUN_BAD: set when a walk's reach was capped (pairs missed)
    float4* movers;   // (build position, slot bits) of the particles that moved more than skin/2 since the build
    uint* defer;      // slots whose grid walk a pass left to its deferred kernel (kDeferCap entries)
    uint* defer_n;    // their count (stats slot ST_DEFER)
    int defer_on;     // 0: every particle walks in its own pass
    float cross2[4];  // (skin of level l / 2)^2: the crossings the skin sizing counts
    int maxn;
    float eps;
    float gfac;       // constraint-gradient factor: spiky gradient coefficient / rho0, times m in the mass mode
    float dt, inv_dt;
    float xsph_c;
    float tol;        // adaptive solve: stop once the judged density error is at most tol
    int judge;        // 0: max (rho - rho0)+ / rho0; 1: its mean over the particles
    float lam0_scale; // factor on the first multiplier (scaled mode; 1 otherwise)
    float warm_kappa; // learned mode: the first multiplier stays within [0, warm_kappa * lambda_0] (0: unbounded)
    float first_dmax; // largest displacement of the first correction of a step (0: unbounded)
    // feature normalisation of the learned warm start
    float sd_ref;     // lambda denominator of an interior particle of the rest lattice
    float inv_nref;   // 1 / neighbour count of the rest lattice
    float inv_gref;   // 1 / sqrt(sd_ref)
    float inv_h;
    float inv_spacing;
};

namespace {

constexpr int kTpb = 256;
constexpr int kMaxStencil = 4;  // largest cell radius of a fallback walk
constexpr uint kMoverCap = 4096;  // movers listed per list build; beyond it every pass bounds displacements by the largest
constexpr uint kMoverNear = 128;  // movers one workgroup stages; beyond it its particles bound displacements by the largest
constexpr uint kDeferCap = 8192;   // grid walks one pass leaves to its deferred kernel; beyond it they walk in the pass
constexpr unsigned kDeferBlocks = 128;  // workgroups of a deferred kernel (one wavefront per queued slot at a time)

// Per-step statistics slots (float/uint in one array).
enum Stat {
    ST_ERR_START = 0,        // max (rho - rho0)+ / rho0 at the step's start positions (judge of the previous step)
    ST_ERR_ITER = 1,         // .. 1 + k: at the start of Jacobi iteration k (k < 16)
    ST_DISP2 = 17,           // max squared displacement since the list build
    ST_OVERFLOW = 18,        // particles whose list overflowed
    ST_ITERS = 19,           // iterations executed (learned mode)
    ST_FALLBACK = 20,        // list passes that had to walk the wide stencil
    ST_ERR_FINAL = 21,       // max error after the solve (learned mode: the last check)
    ST_BAD = 22,             // displacement beyond the wide stencil's reach
    ST_MEAN_ABS_START = 23,  // sum |rho - rho0| / rho0 at the step start (divide by N)
    ST_MPOS_START = 24,      // sum (rho - rho0)+ / rho0 at the step start (divide by N)
    ST_MOVERS = 25,          // movers listed since the list build (uint; kept until the next build)
    ST_MOVERS_PUB = 26,      // ST_MOVERS as of the last lambda pass: the movers a correction pass may read
    ST_CROSS = 27,           // .. 27 + l (l < 4): corrections that carried a particle past half the skin of level l
    ST_LIST_CHUNKS = 31,     // 16-byte chunks stored by the step's list build (uint; the lists the passes walk)
    ST_MPOS_ITER = 32,       // .. 32 + k: the same at the start of Jacobi iteration k (k < 16)
    ST_RUN = 48,             // run accumulators from here on, kept across steps:
    ST_RUN_MAX = 48,         //   max over the judged steps of the max error
    ST_RUN_MPOS = 49,        //   max over the judged steps of the mean positive error
    ST_RUN_ITERS = 50,       //   corrections executed (uint)
    ST_RUN_STEPS = 51,       //   judged steps (uint)
    ST_RUN_FAIL = 52,        //   judged steps whose judged error exceeded tol (uint)
    ST_RUN_BAD = 53,         //   1 once any pass since the reset had its fallback reach capped (uint)
    ST_RUN_CHUNKS = 54,      //   list chunks stored by the list builds (float sum)
    ST_DEFER = 55,           // deferred walks queued by the running pass (uint; zero between passes)
    ST_DEFER_DONE = 56,      // workgroups of the running deferred kernel that finished (uint; zero between passes)
    ST_MOVERS_SORTED = 57,   // leading entries of the mover list already in slot order (uint)
    ST_N = 64
};

// Error of iteration `it` under the solver's judge is within the tolerance.
__device__ inline bool converged(const float* stats, int it, const PbfK& k) {
    return k.judge == 0 ? stats[ST_ERR_ITER + it] <= k.tol : stats[ST_MPOS_ITER + it] <= k.tol * (float)k.n;
}



// 1/sqrt(r2) with the hardware reciprocal square root (v_rsq_f32, within 1 ulp): one instruction in
// place of a correctly rounded square root and division per neighbour pair.
__device__ inline float rinv_of(float r2) { return __builtin_amdgcn_rsqf(r2); }

__device__ inline float pow4(float x) {
    float x2 = x * x;
    return x2 * x2;
}

__device__ inline void atomic_max_pos(float* addr, float v) {
    // non-negative floats order like their bit patterns
    if (v > 0.0f) atomicMax((unsigned int*)addr, __float_as_uint(v));
}

// Block-level max of a non-negative value, then one atomic per block.
__device__ inline void block_max_atomic(float v, float* dst) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    __shared__ float wm[kTpb / 64];
    if (lane == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wm[0];
        for (uint w = 1; w < (blockDim.x >> 6); ++w) m = fmaxf(m, wm[w]);
        atomic_max_pos(dst, m);
    }
    __syncthreads();  // wm is read by thread 0 above: the next reduction of the block must not overwrite it first
}
__device__ inline void block_sum_atomic(float v, float* dst) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    __shared__ float ws[kTpb / 64];
    if (lane == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (uint w = 0; w < (blockDim.x >> 6); ++w) m += ws[w];
        atomicAdd(dst, m);
    }
    __syncthreads();  // ws is read by thread 0 above: the next reduction of the block must not overwrite it first
}

// Block-level sum of a uint, then one atomic per block (and its value added to run_dst as a float).
__device__ inline void block_sum_atomic_u(uint v, uint* dst, float* run_dst) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    __shared__ uint wsu[kTpb / 64];
    if (lane == 0) wsu[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint m = 0;
        for (uint w = 0; w < (blockDim.x >> 6); ++w) m += wsu[w];
        if (m) {
            atomicAdd(dst, m);
            atomicAdd(run_dst, (float)m);
        }
    }
    __syncthreads();
}

__device__ inline uint slot_of(const PbfK& k, bool& valid) {
    uint b = xcd_block(blockIdx.x, gridDim.x);
    uint i = b * blockDim.x + threadIdx.x;
    valid = i < k.n;
    return valid ? i : 0;
}

__device__ inline float3 clamp_wall(float3 x, const SphK& s) {
    x.x = fminf(fmaxf(x.x, s.wall_lo.x), s.wall_hi.x);
    x.y = fminf(fmaxf(x.y, s.wall_lo.y), s.wall_hi.y);
    x.z = fminf(fmaxf(x.z, s.wall_lo.z), s.wall_hi.z);
    return x;
}

// Walks the cells within `rad` of cell c (rows of contiguous slots), fetching four candidates ahead of their use:
// the loads of a group are issued together, so a wavefront keeps several gathers in flight.
// The box of cells [lo, hi] (inclusive, inside the grid).
template <class Fetch, class Body>
__device__ inline void stencil_box(int3 lo, int3 hi, const FGridK& g, const uint* __restrict__ off, Fetch&& fetch, Body&& body) {
    for (int z = lo.z; z <= hi.z; ++z)
        for (int y = lo.y; y <= hi.y; ++y) {
            uint hb = (uint)((z * g.gy + y) * g.gx);
            uint j = off[hb + lo.x], e = off[hb + hi.x + 1];
            for (; j + 3 < e; j += 4) {
                auto a = fetch(j);
                auto b = fetch(j + 1);
                auto cc = fetch(j + 2);
                auto d = fetch(j + 3);
                body(j, a);
                body(j + 1, b);
                body(j + 2, cc);
                body(j + 3, d);
            }
            for (; j < e; ++j) body(j, fetch(j));
        }
}

template <class Fetch, class Body>
__device__ inline void stencil_batched(int3 c, const FGridK& g, const uint* __restrict__ off, Fetch&& fetch, Body&& body, int rad) {
    stencil_box(make_int3(max(c.x - rad, 0), max(c.y - rad, 0), max(c.z - rad, 0)),
                make_int3(min(c.x + rad, g.gx - 1), min(c.y + rad, g.gy - 1), min(c.z + rad, g.gz - 1)), g, off, fetch, body);
}

// List format.  The candidates of slot i come from three z-planes of cells (z - 1, z, z + 1 around its cell); within
// a plane they lie in [b_p, e_p), b_p = the first slot of the plane's first row.  An entry is 16 bits: the plane
// p in bits 14-15 and j - b_p in bits 0-13.  Eight entries make a uint4 chunk; chunk q of slot i is at
// [q * n + i] (a wavefront reads 1 KB contiguous per chunk).  The head of slot i is (b_0, b_1, b_2, count), the
// count with bit 31 set when an entry did not fit (the passes then walk the grid).
constexpr uint kListOverflow = 0x80000000u;
constexpr uint kWide = 0x40000000u;  // list build: an entry's offset did not fit in 14 bits

__device__ inline uint list_slot(uint e, const uint4& hd) {
    // plane 1 sets bit 14, plane 2 bit 15: two selects on bits (an equality chain on p is turned into a
    // table lookup through scratch memory)
    uint b = (e & 0x4000u) ? hd.y : hd.x;
    b = (e & 0x8000u) ? hd.z : b;
    return b + (e & 0x3FFFu);
}
__device__ inline void list_unpack(const uint4& w, const uint4& hd, uint* j) {
    j[0] = list_slot(w.x & 0xFFFFu, hd);
    j[1] = list_slot(w.x >> 16, hd);
    j[2] = list_slot(w.y & 0xFFFFu, hd);
    j[3] = list_slot(w.y >> 16, hd);
    j[4] = list_slot(w.z & 0xFFFFu, hd);
    j[5] = list_slot(w.z >> 16, hd);
    j[6] = list_slot(w.w & 0xFFFFu, hd);
    j[7] = list_slot(w.w >> 16, hd);
}

// The list walk: groups of G chunks (8 G neighbours); the next group's chunks are loaded while the current
// group's neighbours are fetched, and all fetches of a group go out before any of them is used.
template <int G, class Fetch, class Body>
__device__ inline void list_walk(const uint4* __restrict__ l4, const uint4& hd, uint n, Fetch&& fetch, Body&& body) {
    using D = decltype(fetch(0u));
    const uint c = hd.w;
    const uint full = c >> 3;
    uint q = 0;
    if (full) {
        uint4 e[G];
#pragma unroll
        for (int g = 0; g < G; ++g) e[g] = l4[(size_t)min((uint)g, full - 1) * n];
        for (; q + G <= full; q += G) {
            uint4 en[G];
#pragma unroll
            for (int g = 0; g < G; ++g) en[g] = l4[(size_t)min(q + G + g, full - 1) * n];
            uint j[8 * G];
            D d[8 * G];
#pragma unroll
            for (int g = 0; g < G; ++g) list_unpack(e[g], hd, j + 8 * g);
#pragma unroll
            for (int t = 0; t < 8 * G; ++t) d[t] = fetch(j[t]);
#pragma unroll
            for (int t = 0; t < 8 * G; ++t) body(j[t], d[t]);
#pragma unroll
            for (int g = 0; g < G; ++g) e[g] = en[g];
        }
#pragma unroll
        for (int g = 0; g + 1 < G; ++g) {  // fewer than G whole chunks left: e[0 ..] hold them
            if (q + g < full) {
                uint j[8];
                D d[8];
                list_unpack(e[g], hd, j);
#pragma unroll
                for (int t = 0; t < 8; ++t) d[t] = fetch(j[t]);
#pragma unroll
                for (int t = 0; t < 8; ++t) body(j[t], d[t]);
            }
        }
    }
    const uint r = c & 7u;
    if (r) {
        uint j[8];
        D d[8];
        list_unpack(l4[(size_t)full * n], hd, j);
#pragma unroll
        for (int t = 0; t < 7; ++t)
            if ((uint)t < r) d[t] = fetch(j[t]);
#pragma unroll
        for (int t = 0; t < 7; ++t)
            if ((uint)t < r) body(j[t], d[t]);
    }
}

// Squared length with a fixed rounding: the list build and the mover test below both decide "j is in the list
// of i" with it, so they agree bit for bit.
__device__ inline float len2_fixed(float3 r) { return __builtin_fmaf(r.z, r.z, __builtin_fmaf(r.y, r.y, r.x * r.x)); }

// Movers: particles whose displacement since the list build exceeded skin/2 at some correction this step.  The
// correction pass lists each once (build position, slot) the first time it crosses; they stay listed until the
// next build.  Every particle that is not listed is within skin/2 of its build position.  The movers a
// workgroup's particles may meet are staged in LDS with their displacement at the pass's positions: n = -1 when
// the workgroup bounds every displacement by the largest one instead (too many movers in the step, or near the
// workgroup).
struct NearMovers {
    const float4* x;   // build position, slot bits
    const uint* cell;  // build cell hash
    const float* d;    // displacement from the build position at the pass's positions
    int n;
    float dmax;        // n < 0: the largest displacement a neighbour of the workgroup's particles can have
};

// Rounding margin of the reach tests below (their distances are sums of rounded lengths).
constexpr float kReachMargin = 1.0001f;

// Call from every thread of the workgroup (the result is the same for all).  xp: the positions the pass reads.
// A particle i that has moved d_i meets a mover j (moved d_j) only if |xb_i - xb_j| <= h + d_i + d_j, so a mover
// is staged when it lies within h + d_w + d_j of the box of the workgroup's build positions, d_w = the largest
// of skin/2 and the displacements of the workgroup's particles: one fast particle widens only its own
// workgroup's reach and the movers near it.  `published`: read the movers as of the last lambda pass (a
// correction pass appends to the list while it runs).
__device__ inline NearMovers stage_movers(uint i, bool valid, const PbfK& k, float disp2, const float* __restrict__ stats,
                                          const float4* __restrict__ xb, const float4* __restrict__ xp,
                                          const uint* __restrict__ bhash, bool published) {
    __shared__ float4 s_x[kMoverNear];
    __shared__ uint s_cell[kMoverNear];
    __shared__ float s_d[kMoverNear];
    __shared__ float s_box[7][kTpb / 64];
    NearMovers nm{s_x, s_cell, s_d, 0, sqrtf(disp2)};
    if (!(disp2 > k.half_skin2)) return nm;  // no particle has moved more than skin/2: the lists hold every pair
    const uint nmov = __float_as_uint(stats[published ? ST_MOVERS_PUB : ST_MOVERS]);
    if (nmov > kMoverCap) {  // (the list is incomplete: every displacement is bounded by the largest)
        nm.n = -1;
        return nm;
    }
    if (nmov == 0) return nm;
    float b[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
    if (valid) {
        const float4 q = xb[i];
        b[0] = b[3] = q.x, b[1] = b[4] = q.y, b[2] = b[5] = q.z;
        b[6] = len2_fixed(f3(xp[i]) - f3(q));
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            b[c] = fminf(b[c], __shfl_xor(b[c], d, 64));
            b[c + 3] = fmaxf(b[c + 3], __shfl_xor(b[c + 3], d, 64));
        }
        b[6] = fmaxf(b[6], __shfl_xor(b[6], d, 64));
    }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 7; ++c) s_box[c][threadIdx.x >> 6] = b[c];
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        for (int c = 0; c < 3; ++c) {
            b[c] = fminf(b[c], s_box[c][w]);
            b[c + 3] = fmaxf(b[c + 3], s_box[c + 3][w]);
        }
        b[6] = fmaxf(b[6], s_box[6][w]);
    }
    const float dw = fmaxf(0.5f * k.skin, sqrtf(b[6]));
    const float R0 = (k.s.h + dw + sqrtf(disp2)) * kReachMargin;  // every mover's displacement is at most sqrt(disp2)
    // staged in list order (the list is sorted by slot after every correction pass): a prefix over the
    // workgroup per round of blockDim.x movers, so the order in which a particle sums them is the same every run
    __shared__ uint s_cnt[kTpb / 64];
    uint staged = 0;
    float dtake = 0.0f;  // the largest displacement of a mover within reach (this thread's share)
    for (uint m0 = 0; m0 < nmov; m0 += blockDim.x) {
        const uint m = m0 + threadIdx.x;
        bool take = false;
        float4 mv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float dj = 0.0f;
        uint slot = 0;
        if (m < nmov) {
            mv = k.movers[m];
            if (mv.x >= b[0] - R0 && mv.x <= b[3] + R0 && mv.y >= b[1] - R0 && mv.y <= b[4] + R0 && mv.z >= b[2] - R0 && mv.z <= b[5] + R0) {
                slot = min(__float_as_uint(mv.w), k.n - 1u);
                dj = sqrtf(len2_fixed(f3(xp[slot]) - f3(mv)));
                const float R = (k.s.h + dw + dj) * kReachMargin;
                take = mv.x >= b[0] - R && mv.x <= b[3] + R && mv.y >= b[1] - R && mv.y <= b[4] + R && mv.z >= b[2] - R && mv.z <= b[5] + R;
            }
        }
        const uint64_t bal = __ballot(take);
        const uint wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        if (lane == 0) s_cnt[wv] = (uint)__builtin_popcountll(bal);
        __syncthreads();
        uint pos = staged + (uint)__builtin_popcountll(bal & ((1ull << lane) - 1ull)), round = 0;
        for (uint w = 0; w < (blockDim.x >> 6); ++w) {
            pos += w < wv ? s_cnt[w] : 0u;
            round += s_cnt[w];
        }
        if (take && pos < kMoverNear) {
            s_x[pos] = mv;
            s_cell[pos] = bhash[slot];
            s_d[pos] = dj;
        }
        dtake = take ? fmaxf(dtake, dj) : dtake;
        staged += round;
        __syncthreads();  // s_cnt is rewritten by the next round
    }
    nm.n = staged > kMoverNear ? -1 : (int)staged;
    if (nm.n < 0) {
        // too many to stage: a neighbour j of one of the workgroup's particles is either within skin/2 of its
        // build position or a mover within reach, so the largest of those bounds every such d_j
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) dtake = fmaxf(dtake, __shfl_xor(dtake, d, 64));
        if ((threadIdx.x & 63) == 0) s_box[0][threadIdx.x >> 6] = dtake;
        __syncthreads();
        float dm = 0.5f * k.skin;
        for (uint w = 0; w < (blockDim.x >> 6); ++w) dm = fmaxf(dm, s_box[0][w]);
        nm.dmax = dm;
        __syncthreads();  // s_box is reused by the next staging
    }
    return nm;
}

// Keeps the listed movers in slot order (one workgroup), so that every pass reads them in one order whatever
// order the correction pass's atomics listed them in: the sums over movers, and with them the positions, are the
// same from run to run.  Runs after every correction pass.  The entries up to stats[ST_MOVERS_SORTED] are in
// order from the previous run; the new tail's keys (slot, index) are sorted in LDS (bitonic), then both runs
// merge by rank: an entry's place is its index in its own run plus the number of keys below it in the other.
constexpr uint kSortThreads = 1024;
__global__ __launch_bounds__(kSortThreads) void k_sort_movers(float4* __restrict__ movers, float* __restrict__ stats) {
    __shared__ float4 s_out[kMoverCap];
    __shared__ uint2 s_tail[kMoverCap];  // (slot, index) of the new entries
    __shared__ uint s_head[kMoverCap];   // slots of the sorted entries
    const uint n = __float_as_uint(stats[ST_MOVERS]);
    const uint h = min(__float_as_uint(stats[ST_MOVERS_SORTED]), n);
    if (n > kMoverCap || h == n) return;  // (beyond the cap the passes do not read the list)
    const uint tn = n - h;
    uint np = 1u;
    while (np < tn) np <<= 1;
    for (uint t = threadIdx.x; t < np; t += blockDim.x)
        s_tail[t] = t < tn ? make_uint2(__float_as_uint(movers[h + t].w), h + t) : make_uint2(0xFFFFFFFFu, 0u);
    for (uint t = threadIdx.x; t < h; t += blockDim.x) s_head[t] = __float_as_uint(movers[t].w);
    __syncthreads();
    for (uint kk = 2u; kk <= np; kk <<= 1)
        for (uint jj = kk >> 1; jj > 0u; jj >>= 1) {
            for (uint t = threadIdx.x; t < np; t += blockDim.x) {
                const uint u = t ^ jj;
                if (u > t) {
                    const uint2 a = s_tail[t], c = s_tail[u];
                    if ((a.x > c.x) == ((t & kk) == 0u)) {
                        s_tail[t] = c;
                        s_tail[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    // slots are distinct (a particle is listed once), so the ranks place every entry exactly once
    for (uint t = threadIdx.x; t < h; t += blockDim.x) {
        const uint key = s_head[t];
        uint lo = 0u, hi = tn;  // tail keys below key
        while (lo < hi) {
            const uint mid = (lo + hi) >> 1;
            if (s_tail[mid].x < key) lo = mid + 1u;
            else hi = mid;
        }
        s_out[t + lo] = movers[t];
    }
    for (uint t = threadIdx.x; t < tn; t += blockDim.x) {
        const uint key = s_tail[t].x;
        uint lo = 0u, hi = h;  // head keys below key
        while (lo < hi) {
            const uint mid = (lo + hi) >> 1;
            if (s_head[mid] < key) lo = mid + 1u;
            else hi = mid;
        }
        s_out[t + lo] = movers[s_tail[t].y];
    }
    __syncthreads();
    for (uint t = threadIdx.x; t < n; t += blockDim.x) movers[t] = s_out[t];
    if (threadIdx.x == 0) stats[ST_MOVERS_SORTED] = __uint_as_float(n);
}

// Neighbour candidates of slot i for a pass that needs every j with |x_i - x_j| <= h at the current positions;
// `fetch(j)` loads what the pass needs of j, `body(j, data)` consumes it.  `p` is the current position of i,
// `xb` the build positions, d = displacement since the build, D = its largest value (disp2 = D^2).
// mode 1: the 3x3x3 stencil of a grid built on the current positions (cell >= h).
// mode 0: the skin list, which holds every j with |xb_i - xb_j| <= h + skin.  A pair within h now was within
// h + d_i + d_j at the build.  With movers staged (nm.n >= 0): every particle that is not a mover has d <= skin/2,
// so a particle with d_i <= skin/2 walks its list, which holds every such j, and then the staged movers its list
// does not hold (the build's own test on the build positions and cells decides that, so no pair is visited
// twice); a mover or an overflowed list walks the build-time grid over the cells within L = h + d_i + min(skin/2,
// D) of its build position, which hold every j that is not a mover, and then the staged movers outside those
// cells.
// Without staged movers (nm.n < 0) a particle walks its list while d_i + D_w <= skin, else the grid within
// L = h + d_i + D_w, where D_w = nm.dmax bounds the displacement of any neighbour: the largest of skin/2 and the
// displacements of the movers within the workgroup's reach (too many to stage), or D when the mover list is
// incomplete.  The grid radius is capped at kMaxStencil cells (beyond it the step and the run are flagged).
// With k.defer_on and movers staged, the grid walk of a mover is queued instead (deferred walks, below): the
// function then returns false and the pass leaves slot i to its deferred kernel.
template <int G = 1, class Fetch, class Body>
__device__ inline bool for_neighbors(uint i, int mode, const PbfK& k, const uint* __restrict__ list, const uint4* __restrict__ cnt,
                                     const uint* __restrict__ bhash, const uint* __restrict__ off, float disp2, float3 p,
                                     const float4* __restrict__ xb, const NearMovers& nm, Fetch&& fetch, Body&& body) {
    if (mode == 1) {
        stencil_batched(fg_unhash(bhash[i], k.g), k.g, off, fetch, body, 1);
        return true;
    }
    const uint4 hd = cnt[i];
    bool ok = hd.w <= (uint)k.maxn;
    bool movers = false, mover = false;
    float di = 0.0f;
    if (disp2 > k.half_skin2) {
        const float di2 = len2_fixed(p - f3(xb[i]));
        di = sqrtf(di2);
        if (nm.n >= 0) {
            mover = di2 > k.half_skin2;
            ok = ok && !mover;
            movers = nm.n > 0;
        } else {
            ok = ok && di + nm.dmax <= k.skin;
        }
    }
    if (ok) {
        list_walk<G>((const uint4*)list + i, hd, k.n, fetch, body);
        if (movers) {
            const float3 bi = f3(xb[i]);
            const int3 ci = fg_unhash(bhash[i], k.g);
            for (int m = 0; m < nm.n; ++m) {
                const float4 mv = nm.x[m];
                const float r2 = len2_fixed(bi - f3(mv));
                const float R = (k.s.h + di + nm.d[m]) * kReachMargin;
                if (r2 > R * R) continue;  // out of reach now
                const int3 cj = fg_unhash(nm.cell[m], k.g);
                const bool listed = abs(cj.x - ci.x) <= 1 && abs(cj.y - ci.y) <= 1 && abs(cj.z - ci.z) <= 1 && r2 <= k.rc2;
                if (!listed) {
                    const uint j = __float_as_uint(mv.w);
                    body(j, fetch(j));
                }
            }
        }
    } else {
        // every j within L of i's build position: the cells that the box [x_b - L, x_b + L] touches (3 or 4 per
        // axis once L exceeds a cell, not the 5 of a cube of radius 2 around the cell)
        const uint64_t wave = __ballot(1);
        if (__lane_id() == (uint)__builtin_ctzll(wave)) atomicAdd(k.fallback, (uint)__builtin_popcountll(wave));
        const bool staged = nm.n >= 0;
        if (mover && k.defer_on) {  // (movers are at most kMoverCap <= kDeferCap: the queue never fills)
            const uint t = atomicAdd(k.defer_n, 1u);
            if (t < kDeferCap) {
                k.defer[t] = i;
                return false;
            }
        }
        const float D = sqrtf(disp2);
        const float Lwant = (k.s.h + di + (staged ? fminf(0.5f * k.skin, D) : nm.dmax)) * kReachMargin;
        const float L = fminf(Lwant, kMaxStencil * k.g.cell);
        if (Lwant > L) {  // the capped walk may miss pairs: flag the step and the run (never silently)
            atomicOr(k.bad, 1u);
            atomicOr(k.bad + (ST_RUN_BAD - ST_BAD), 1u);
        }
        const float3 c = f3(xb[i]);
        const int3 lo = fg_pos(c - f3(L, L, L), k.g), hi = fg_pos(c + f3(L, L, L), k.g);
        stencil_box(lo, hi, k.g, off, fetch, body);
        for (int m = 0; m < (staged ? nm.n : 0); ++m) {  // the movers beyond those cells
            const float4 mv = nm.x[m];
            const int3 cj = fg_unhash(nm.cell[m], k.g);
            if (cj.x >= lo.x && cj.x <= hi.x && cj.y >= lo.y && cj.y <= hi.y && cj.z >= lo.z && cj.z <= hi.z) continue;
            const float R = (k.s.h + di + nm.d[m]) * kReachMargin;
            if (len2_fixed(c - f3(mv)) > R * R) continue;
            const uint j = __float_as_uint(mv.w);
            body(j, fetch(j));
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------ deferred walks
// A mover that a pass would send to the grid while its movers are staged is queued instead, and after the pass
// one wavefront walks it (the k_*_deferred kernels): lane l takes every 64th
// candidate of the build-time grid cells within L = h + d_i + min(skin/2, D) of its build position, then every
// 64th listed mover outside those cells (within h + d_i + D), and the lanes' partial sums are added across the
// wavefront.  A few such walks of several hundred candidates then no longer hold whole wavefronts of 40-60-entry
// list walks.  The queue count is stats slot ST_DEFER; the last workgroup of the deferred kernel clears it.
// Which particles are queued depends on the positions alone (movers, at most kMoverCap), and each is summed by
// a fixed split across the lanes, so the results do not depend on the order the queue was filled in.
__device__ inline float wave_sum(float v) {  // every lane ends with the same sum (each step adds a symmetric pair)
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
template <bool kCoop>
__device__ inline float wsum(float v) {
    if constexpr (kCoop) return wave_sum(v);
    return v;
}
template <bool kCoop>
__device__ inline float3 wsum3(float3 v) {
    if constexpr (kCoop) return make_float3(wave_sum(v.x), wave_sum(v.y), wave_sum(v.z));
    return v;
}
template <bool kCoop>
__device__ inline bool writer() {  // the lane that stores a slot's results
    if constexpr (kCoop) return __lane_id() == 0;
    return true;
}
__device__ inline uint wave_index() { return (blockIdx.x * blockDim.x + threadIdx.x) >> 6; }
__device__ inline uint wave_count() { return (gridDim.x * blockDim.x) >> 6; }
__device__ inline uint defer_count(const float* __restrict__ stats) { return min(__float_as_uint(stats[ST_DEFER]), kDeferCap); }
// Every thread, at the end of a deferred kernel: the last workgroup to finish clears the queue for the next pass
// (every workgroup read the count before it counted itself done).
__device__ inline void defer_done(float* __restrict__ stats) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd((uint*)&stats[ST_DEFER_DONE], 1u) == gridDim.x - 1u) {
            atomicExch((uint*)&stats[ST_DEFER], 0u);
            atomicExch((uint*)&stats[ST_DEFER_DONE], 0u);
        }
    }
}

template <class Fetch, class Body>
__device__ inline void coop_neighbors(uint i, float di, const PbfK& k, const uint* __restrict__ bhash, const uint* __restrict__ off,
                                      float disp2, const float4* __restrict__ xb, uint nmov, Fetch&& fetch, Body&& body) {
    const uint lane = __lane_id();
    const float D = sqrtf(disp2);
    const float Lwant = (k.s.h + di + fminf(0.5f * k.skin, D)) * kReachMargin;
    const float L = fminf(Lwant, kMaxStencil * k.g.cell);
    if (Lwant > L && lane == 0) {
        atomicOr(k.bad, 1u);
        atomicOr(k.bad + (ST_RUN_BAD - ST_BAD), 1u);
    }
    const float3 c = f3(xb[i]);
    const int3 lo = fg_pos(c - f3(L, L, L), k.g), hi = fg_pos(c + f3(L, L, L), k.g);
    // the box's rows of contiguous slots, four at a time: their bounds, then one candidate per lane of each, are
    // loaded together (rows are mostly shorter than a wavefront, so a row at a time waits on every load)
    const int ny = hi.y - lo.y + 1, nrow = ny * (hi.z - lo.z + 1);
    using Dt = decltype(fetch(0u));
    for (int r0 = 0; r0 < nrow; r0 += 4) {
        uint b[4], e[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            b[u] = e[u] = 0u;
            if (r0 + u < nrow) {
                const int r = r0 + u, z = lo.z + r / ny, y = lo.y + r % ny;
                const uint hb = (uint)((z * k.g.gy + y) * k.g.gx);
                b[u] = off[hb + lo.x];
                e[u] = off[hb + hi.x + 1];
            }
        }
        Dt dd[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b[u] + lane < e[u]) dd[u] = fetch(b[u] + lane);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b[u] + lane < e[u]) body(b[u] + lane, dd[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)  // rows longer than a wavefront
            for (uint j = b[u] + lane + 64u; j < e[u]; j += 64u) body(j, fetch(j));
    }
    // the listed movers outside the box, four loads per lane at a time
    const float R = (k.s.h + di + D) * kReachMargin;
    const uint nm = min(nmov, kMoverCap);
    for (uint m0 = lane; m0 < nm; m0 += 256u) {
        float4 mv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) mv[u] = m0 + 64u * u < nm ? k.movers[m0 + 64u * u] : make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!(len2_fixed(c - f3(mv[u])) <= R * R)) continue;
            const uint j = min(__float_as_uint(mv[u].w), k.n - 1u);
            const int3 cj = fg_unhash(bhash[j], k.g);
            if (cj.x >= lo.x && cj.x <= hi.x && cj.y >= lo.y && cj.y <= hi.y && cj.z >= lo.z && cj.z <= hi.z) continue;
            body(j, fetch(j));
        }
    }
}

struct XV {  // position and velocity of a neighbour
    float4 x, v;
};

// ------------------------------------------------------------------------------------------ predict
// rxPBDSPH::calDensity + calForceExtAndVisc + integrate (rx_sph_pbd.cpp:599-689, 1409-1446), fused.
// Reads x, v (w unused); writes (x*, rho_start) to xs.  Also judges the previous step: max (rho - rho0)+ / rho0.
// One slot: walk(p, fetch, body) visits the neighbours (false: deferred); e = (rho - rho0) / rho0.
template <bool kCoop, class Walk>
__device__ inline bool predict_slot(uint i, const float4* __restrict__ x, const float4* __restrict__ v, float4* __restrict__ xs,
                                    float* __restrict__ dens, const PbfK& k, Walk&& walk, float& e) {
    float3 p = f3(x[i]), vi = f3(v[i]);
    float d = 0.0f;
    float3 fv = f3(0, 0, 0);
    if (!walk(p, [&](uint j) { return x[j]; },
              [&](uint j, const float4& xj) {
                  float3 r = p - f3(xj);
                  float r2 = dot3(r, r);
                  if (r2 <= k.s.h2) {
                      float qq = k.s.h2 - r2;
                      d += qq * qq * qq;
                      if (r2 > 0.0f) fv += (k.s.h - r2 * rinv_of(r2)) * (f3(v[j]) - vi);  // velocity read within h only
                  }
              }))
        return false;
    d = wsum<kCoop>(d);
    fv = wsum3<kCoop>(fv);
    d *= k.s.mass * k.s.wpoly6;
    float3 a = (k.s.visc * k.s.mass * k.s.lwvisc / d) * fv + k.s.gravity;  // viscosity / rho_i (rx_sph_pbd.cpp:676)
    float3 vn = vi + k.dt * a;
    float3 xn = clamp_wall(p + k.dt * vn, k.s);
    if (writer<kCoop>()) {
        xs[i] = make_float4(xn.x, xn.y, xn.z, d);  // w: density at the step start (features)
        dens[i] = d;
    }
    e = (d - k.s.rho0) * k.s.inv_rho0;
    return true;
}

__global__ __launch_bounds__(kTpb) void k_predict(const float4* __restrict__ x, const float4* __restrict__ v, float4* __restrict__ xs,
                                                  float* __restrict__ dens, float* __restrict__ stats, int mode, PbfK k,
                                                  const uint* __restrict__ list, const uint4* __restrict__ cnt,
                                                  const uint* __restrict__ bhash, const uint* __restrict__ off,
                                                  const float4* __restrict__ xb) {
    bool valid;
    uint i = slot_of(k, valid);
    float disp2 = stats[ST_DISP2];
    const NearMovers nm = mode == 0 ? stage_movers(i, valid, k, disp2, stats, xb, x, bhash, false) : NearMovers{nullptr, nullptr, nullptr, 0, 0.0f};
    float err = 0.0f, aerr = 0.0f;
    float e;
    if (valid && predict_slot<false>(i, x, v, xs, dens, k,
                                     [&](float3 p, auto&& fetch, auto&& body) {
                                         return for_neighbors<1>(i, mode, k, list, cnt, bhash, off, disp2, p, xb, nm, fetch, body);
                                     },
                                     e)) {
        err = fmaxf(e, 0.0f);
        aerr = fabsf(e);
    }
    block_max_atomic(err, &stats[ST_ERR_START]);
    block_sum_atomic(aerr, &stats[ST_MEAN_ABS_START]);
    block_sum_atomic(err, &stats[ST_MPOS_START]);
}

__global__ __launch_bounds__(kTpb) void k_predict_deferred(const float4* __restrict__ x, const float4* __restrict__ v,
                                                           float4* __restrict__ xs, float* __restrict__ dens,
                                                           float* __restrict__ stats, PbfK k, const uint* __restrict__ bhash,
                                                           const uint* __restrict__ off, const float4* __restrict__ xb) {
    const uint nq = defer_count(stats), nmov = __float_as_uint(stats[ST_MOVERS]);
    const float disp2 = stats[ST_DISP2];
    float err = 0.0f, aerr = 0.0f, perr = 0.0f;
    for (uint q = wave_index(); q < nq; q += wave_count()) {
        const uint i = k.defer[q];
        const float di = sqrtf(len2_fixed(f3(x[i]) - f3(xb[i])));
        float e;
        predict_slot<true>(i, x, v, xs, dens, k,
                           [&](float3, auto&& fetch, auto&& body) {
                               coop_neighbors(i, di, k, bhash, off, disp2, xb, nmov, fetch, body);
                               return true;
                           },
                           e);
        if (writer<true>()) {
            err = fmaxf(err, fmaxf(e, 0.0f));
            perr += fmaxf(e, 0.0f);
            aerr += fabsf(e);
        }
    }
    block_max_atomic(err, &stats[ST_ERR_START]);
    block_sum_atomic(aerr, &stats[ST_MEAN_ABS_START]);
    block_sum_atomic(perr, &stats[ST_MPOS_START]);
    defer_done(stats);
}

// Zeroes the statistics slots whose bit is set in `mask` (64 threads).
__global__ void k_zero_slots(float* __restrict__ stats, unsigned long long mask) {
    const uint t = threadIdx.x;
    if (t < (uint)ST_N && ((mask >> t) & 1ull)) stats[t] = 0.0f;
}

// Folds the judge of the previous step (its final positions, measured by this step's prediction pass) into
// the run accumulators.
__global__ void k_fold_run(float* __restrict__ stats, PbfK k) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float mx = stats[ST_ERR_START], mp = stats[ST_MPOS_START] / (float)k.n;
    stats[ST_RUN_MAX] = fmaxf(stats[ST_RUN_MAX], mx);
    stats[ST_RUN_MPOS] = fmaxf(stats[ST_RUN_MPOS], mp);
    uint* u = (uint*)stats;
    u[ST_RUN_STEPS] += 1u;
    if ((k.judge == 0 ? mx : mp) > k.tol) u[ST_RUN_FAIL] += 1u;
}

// ------------------------------------------------------------------------------------------ list + lambda
// Builds the skin list of slot i from the new grid (cell = h + skin) and computes lambda at x*
// (rx_sph_pbd.cpp:699-822: C = rho/rho0 - 1, sd with the self term).  Writes xl = (x*, lambda).  When feat is
// set it also writes the 16 features of the learned warm start (12 used, 4 zero padding) for slot i.
// `cands(plane, f)` visits the 27 cells around slot i as 9 rows of contiguous slots in z, y order: for each
// z-plane pl (0, 1, 2 = z - 1, z, z + 1) inside the grid it calls plane(pl, base, span) with the first slot of the
// plane's first row and the slots up to the end of its last row, then f(j, xs[j], base - (pl << 14)) for the
// slots j of its rows in ascending order.
// xrow (may be null): the features are also written, standardised as bf16, to this LDS row of the network
// staged in snet (the fused inference).
template <class Cands>
__device__ inline void list_lambda_slot(uint i, float4 p4, Cands&& cands, const float4* __restrict__ xs, float4* __restrict__ xl,
                                        float4* __restrict__ xbuild,
                                        float* __restrict__ dens, uint* __restrict__ list, uint4* __restrict__ cnt,
                                        float* __restrict__ stats, const PbfK& k, float* __restrict__ feat,
                                        const float4* __restrict__ v, float& emax, float& esum, uint& chunks,
                                        unsigned short* xrow = nullptr, const MlpShared* snet = nullptr) {
    const bool want_feat = feat || xrow;
    float3 p = f3(p4);
    float d = 0.0f, sd = 0.0f;
    float3 gs = f3(0, 0, 0);
    uint c = 0;
    // plane bases for the list head; a plane spanning more than 2^14 slots sets kWide in c (no entries are
    // stored and the passes walk the grid)
    uint base0 = 0u, base1 = 0u, base2 = 0u;
    auto plane = [&](uint pl, uint base, uint span) {
        base0 = pl == 0u ? base : base0;
        base1 = pl == 1u ? base : base1;
        base2 = pl == 2u ? base : base2;
        c = span > 0x4000u ? (c | kWide) : c;
    };
    // learned-start extras
    float nn = 0.0f, wsum = 0.0f, wcs = 0.0f, ring = 0.0f, nring = 0.0f;
    const float h2q = 0.25f * k.s.h2;
    uint4* l4 = (uint4*)list + i;  // chunk q of slot i at l4[q * n]
    uint b0 = 0, b1 = 0, b2 = 0, b3 = 0;  // the open chunk as a 128-bit shift register, newest entry at the top
    cands(plane, [&](uint j, float4 q4, uint pl_base) {  // q4 = (x*_j, density at the step start)
        float3 r = p - f3(q4);
        float r2 = len2_fixed(r);
        if (r2 <= k.rc2) {
            if (c < (uint)k.maxn) {
                const uint e = j - pl_base;  // pl_base = base - (plane << 14): the entry's plane bits come with it
                b0 = (b0 >> 16) | (b1 << 16);
                b1 = (b1 >> 16) | (b2 << 16);
                b2 = (b2 >> 16) | (b3 << 16);
                b3 = (b3 >> 16) | (e << 16);
                if ((c & 7u) == 7u) l4[(size_t)(c >> 3) * k.n] = make_uint4(b0, b1, b2, b3);
            }
            ++c;
            if (r2 <= k.s.h2) {
                float q2 = k.s.h2 - r2;
                float w = q2 * q2 * q2;
                d += w;
                float3 g = f3(0, 0, 0);
                if (r2 > 0.0f) {
                    float ri = rinv_of(r2);
                    float q = k.s.h - r2 * ri;
                    g = (k.gfac * q * q * ri) * r;
                    gs += g;
                    sd += dot3(g, g);
                }
                if (want_feat) {
                    float cj = q4.w * k.s.inv_rho0 - 1.0f;
                    nn += 1.0f;
                    wsum += w;
                    wcs += w * cj;
                    if (r2 > h2q) {
                        ring += cj;
                        nring += 1.0f;
                    }
                }
            }
        }
    });
    if (c <= (uint)k.maxn && (c & 7u)) {  // the last, partial chunk: shift its r entries to the bottom
        const uint r = c & 7u;
#pragma unroll
        for (uint t = 0; t < 7u; ++t) {
            if (t < 8u - r) {
                b0 = (b0 >> 16) | (b1 << 16);
                b1 = (b1 >> 16) | (b2 << 16);
                b2 = (b2 >> 16) | (b3 << 16);
                b3 = b3 >> 16;
            }
        }
        l4[(size_t)(c >> 3) * k.n] = make_uint4(b0, b1, b2, b3);
    }
    const bool over = c > (uint)k.maxn;  // (kWide makes c larger than any capacity)
    cnt[i] = make_uint4(base0, base1, base2, over ? (c | kListOverflow) : c);
    chunks += over ? 0u : (c + 7u) >> 3;
    if (over) atomicAdd((unsigned int*)&stats[ST_OVERFLOW], 1u);
    d *= k.s.mass * k.s.wpoly6;
    float sgs = dot3(gs, gs);
    sd += sgs;
    float C = d * k.s.inv_rho0 - 1.0f;
    float lam = -C / (sd + k.eps);
    xl[i] = make_float4(p.x, p.y, p.z, lam * k.lam0_scale);
    xbuild[i] = make_float4(p.x, p.y, p.z, 0.0f);  // w: mover flag, set by the correction pass
    dens[i] = d;
    float err = fmaxf(C, 0.0f);
    emax = fmaxf(emax, err);
    esum += err;
    if (want_feat) {
        float wall = fminf(fminf(fminf(p.x - k.s.wall_lo.x, k.s.wall_hi.x - p.x), fminf(p.y - k.s.wall_lo.y, k.s.wall_hi.y - p.y)),
                           fminf(p.z - k.s.wall_lo.z, k.s.wall_hi.z - p.z));
        float f[12];
        f[0] = C;
        f[1] = lam * k.sd_ref;
        f[2] = sd / k.sd_ref - 1.0f;
        f[3] = nn * k.inv_nref - 1.0f;
        f[4] = sqrtf(sgs) * k.inv_gref;
        f[5] = gs.y * k.inv_gref;
        f[6] = wsum > 0.0f ? wcs / wsum : 0.0f;       // W-weighted mean of the step-start C
        f[7] = nring > 0.0f ? ring / nring : 0.0f;     // mean step-start C of the outer ring (h/2 < r <= h)
        f[8] = p4.w * k.s.inv_rho0 - 1.0f;
        f[9] = (p4.w - d) / d;  // div v dt from the continuity equation: -(rho* - rho_start) / rho*
        f[10] = fminf(fmaxf(wall * k.inv_h, 0.0f), 1.0f) - 0.5f;
        f[11] = v[i].y * k.dt * k.inv_spacing;
        if (feat) {
            float4* fo = (float4*)(feat + (size_t)i * 16);
            fo[0] = make_float4(f[0], f[1], f[2], f[3]);
            fo[1] = make_float4(f[4], f[5], f[6], f[7]);
            fo[2] = make_float4(f[8], f[9], f[10], f[11]);
            fo[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (xrow) mlp_write_row(f, xrow, *snet);
    }
}

// The candidates of the list build (see list_lambda_slot for the contract).
template <class P, class F>
__device__ inline void build_candidates(uint i, const PbfK& k, const float4* __restrict__ xs, const uint* __restrict__ bhash,
                                        const uint* __restrict__ off, P&& plane, F&& f) {
    const FGridK& g = k.g;
    const int3 c = fg_unhash(bhash[i], g);
    const int x0 = max(c.x - 1, 0), x1 = min(c.x + 1, g.gx - 1), y0 = max(c.y - 1, 0), y1 = min(c.y + 1, g.gy - 1);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = c.z + dz;
        if (z < 0 || z >= g.gz) continue;
        const uint pl = (uint)(dz + 1);
        const uint zb = (uint)(z * g.gy);
        const uint base = off[(zb + (uint)y0) * g.gx + x0];
        plane(pl, base, off[(zb + (uint)y1) * g.gx + x1 + 1] - base);
        const uint pb = base - (pl << 14);
        for (int y = y0; y <= y1; ++y) {
            const uint hb = (zb + (uint)y) * g.gx;
            uint j = off[hb + x0];
            const uint e = off[hb + x1 + 1];
            for (; j + 3 < e; j += 4) {
                const float4 a = xs[j], b = xs[j + 1], cc = xs[j + 2], dd = xs[j + 3];
                f(j, a, pb);
                f(j + 1, b, pb);
                f(j + 2, cc, pb);
                f(j + 3, dd, pb);
            }
            for (; j < e; ++j) f(j, xs[j], pb);
        }
    }
}

// One thread per slot, in slot order (the default): the 27-cell walk reads candidates from global memory; the
// state is in cell order, so a wavefront's candidates are few distinct cells and mostly L1/L2 hits.
// kNet (learned warm start): the warm-start network runs on the features while they are in registers (bf16 MFMA,
// the network staged in LDS) and its multiplier, times w_scale, replaces lambda_0 as the first multiplier.
// Held to 64 registers (eight wavefronts per SIMD): the candidate loads are latency-bound, and at 1M particles
// the learned pass runs 5% faster than at 80 registers, the plain one 3%.
template <bool kNet>
__global__ __launch_bounds__(kTpb) __attribute__((amdgpu_waves_per_eu(8))) void k_list_lambda_slots(const float4* __restrict__ xs, float4* __restrict__ xl,
                                                            float4* __restrict__ xbuild, float* __restrict__ dens, uint* __restrict__ list,
                                                            uint4* __restrict__ cnt, float* __restrict__ stats, PbfK k,
                                                            const uint* __restrict__ bhash, const uint* __restrict__ off,
                                                            float* __restrict__ feat, const float4* __restrict__ v, MlpDev net,
                                                            float w_scale) {
    __shared__ MlpShared s_net[1];
    __shared__ __align__(16) unsigned short s_rows[kNet ? kTpb * kMlpIn : 1];
    if (kNet) {
        mlp_stage(net, s_net[0]);
        __syncthreads();
    }
    bool valid;
    uint i = slot_of(k, valid);
    float emax = 0.0f, esum = 0.0f;
    uint chunks = 0;
    unsigned short* xrow = kNet ? s_rows + threadIdx.x * kMlpIn : nullptr;
    if (valid) {
        const float4 p4 = xs[i];
        auto cands = [&](auto&& plane, auto&& f) { build_candidates(i, k, xs, bhash, off, plane, f); };
        list_lambda_slot(i, p4, cands, xs, xl, xbuild, dens, list, cnt, stats, k, feat, v, emax, esum, chunks, xrow, s_net);
    }
    if (kNet) {  // every lane of the wavefront takes part (rows of lanes past the end are zero; their outputs unused)
        if (!valid) ((uint4*)xrow)[0] = ((uint4*)xrow)[1] = make_uint4(0u, 0u, 0u, 0u);
        const float y = mlp_predict_rows(s_rows + (threadIdx.x & ~63u) * kMlpIn, s_net[0], net);
        if (valid) xl[i].w = warm_clamp(y * w_scale, xl[i].w, k.warm_kappa);  // xl.w held lambda_0
    }
    block_max_atomic(emax, &stats[ST_ERR_ITER + 0]);
    block_sum_atomic(esum, &stats[ST_MPOS_ITER + 0]);
    block_sum_atomic_u(chunks, (uint*)&stats[ST_LIST_CHUNKS], &stats[ST_RUN_CHUNKS]);
}

// Alternative with LDS staging (RX_PBF_OPT_LIST_TILE 1..4).  One workgroup per segment of kSegW cells along x
// of one (y, z) row.  The candidates of the whole segment
// are the slots of 9 rows (y - 1 .. y + 1, z - 1 .. z + 1) over the cells x0 - 1 .. x1 + 1: nine contiguous
// slot ranges, staged into LDS with coalesced loads.  Every home slot then walks its 27 cells from LDS.  A
// segment whose neighbourhood exceeds the LDS tile (very dense regions) reads the candidates from global
// memory instead; both paths visit the same candidates in the same order.
template <int kSegW>
inline unsigned list_segments(const FGridK& g) { return (unsigned)(((g.gx + kSegW - 1) / kSegW) * g.gy * g.gz); }

// kSegW cells per segment, kTileCap staged float4 entries, kThreads threads per block (a multiple of 64, <= 256).
template <int kSegW, int kTileCap, int kThreads>
__global__ __launch_bounds__(kThreads) void k_list_lambda(const float4* __restrict__ xs, float4* __restrict__ xl, float4* __restrict__ xbuild,
                                                      float* __restrict__ dens, uint* __restrict__ list, uint4* __restrict__ cnt,
                                                      float* __restrict__ stats, PbfK k, const uint* __restrict__ bhash,
                                                      const uint* __restrict__ off, float* __restrict__ feat,
                                                      const float4* __restrict__ v) {
    __shared__ float4 tile[kTileCap > 0 ? kTileCap : 1];
    __shared__ uint s_lds0[10], s_glob0[9], s_len[9], s_valid[9];
    __shared__ uint s_coff[9][kSegW + 3];
    const FGridK& g = k.g;
    const int nsx = (g.gx + kSegW - 1) / kSegW;
    const uint seg = xcd_block(blockIdx.x, gridDim.x);
    const int sx = (int)(seg % (uint)nsx);
    const uint yz = seg / (uint)nsx;
    const int y = (int)(yz % (uint)g.gy), z = (int)(yz / (uint)g.gy);
    const int x0 = sx * kSegW, x1 = min(x0 + kSegW, g.gx) - 1;
    const uint hb = (uint)((z * g.gy + y) * g.gx);
    const uint hs = off[hb + x0], he = off[hb + x1 + 1];
    if (hs == he) return;  // the same for every thread of the block
    const int xl0 = max(x0 - 1, 0), xh0 = min(x1 + 1, g.gx - 1), ncol = xh0 - xl0 + 1;
    if (threadIdx.x < 9) {
        const int r = threadIdx.x, zz = z + r / 3 - 1, yy = y + r % 3 - 1;
        uint g0 = 0, len = 0;
        s_valid[r] = (zz >= 0 && zz < g.gz && yy >= 0 && yy < g.gy) ? 1u : 0u;
        if (zz >= 0 && zz < g.gz && yy >= 0 && yy < g.gy) {
            const uint rb = (uint)((zz * g.gy + yy) * g.gx);
            g0 = off[rb + xl0];
            len = off[rb + xh0 + 1] - g0;
            for (int t = 0; t <= ncol; ++t) s_coff[r][t] = off[rb + xl0 + t] - g0;
        } else {
            for (int t = 0; t <= ncol; ++t) s_coff[r][t] = 0;
        }
        s_glob0[r] = g0;
        s_len[r] = len;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint a = 0;
        for (int r = 0; r < 9; ++r) {
            s_lds0[r] = a;
            a += s_len[r];
        }
        s_lds0[9] = a;
    }
    __syncthreads();
    const bool tiled = s_lds0[9] <= (uint)kTileCap;
    if (tiled) {
        const uint total = s_lds0[9];
        for (uint t = threadIdx.x; t < total; t += blockDim.x) {
            int r = 0;
            while (t >= s_lds0[r + 1]) ++r;
            tile[t] = xs[s_glob0[r] + (t - s_lds0[r])];
        }
        __syncthreads();
    }
    float emax = 0.0f, esum = 0.0f;
    uint chunks = 0;
    for (uint i = hs + threadIdx.x; i < he; i += blockDim.x) {
        const int cx = (int)(bhash[i] % (uint)g.gx);
        const int a = max(cx - 1, 0) - xl0, b = min(cx + 1, g.gx - 1) - xl0 + 1;
        const float4 p4 = tiled ? tile[s_lds0[4] + (i - s_glob0[4])] : xs[i];
        auto cands = [&](auto&& plane, auto&& f) {
            for (int t = 0; t < 3; ++t) {
                int rf = -1, rl = -1;
                for (int r = 3 * t; r < 3 * t + 3; ++r)
                    if (s_valid[r]) {
                        if (rf < 0) rf = r;
                        rl = r;
                    }
                if (rf < 0) continue;
                const uint base = s_glob0[rf] + s_coff[rf][a];
                plane((uint)t, base, s_glob0[rl] + s_coff[rl][b] - base);
                const uint pb = base - ((uint)t << 14);
                for (int r = rf; r <= rl; ++r) {
                    const uint cs = s_coff[r][a], ce = s_coff[r][b], gb = s_glob0[r];
                    if (tiled) {
                        const uint lb = s_lds0[r];
                        for (uint u = cs; u < ce; ++u) f(gb + u, tile[lb + u], pb);
                    } else {
                        for (uint u = cs; u < ce; ++u) f(gb + u, xs[gb + u], pb);
                    }
                }
            }
        };
        list_lambda_slot(i, p4, cands, xs, xl, xbuild, dens, list, cnt, stats, k, feat, v, emax, esum, chunks);
    }
    block_max_atomic(emax, &stats[ST_ERR_ITER + 0]);
    block_sum_atomic(esum, &stats[ST_MPOS_ITER + 0]);
    block_sum_atomic_u(chunks, (uint*)&stats[ST_LIST_CHUNKS], &stats[ST_RUN_CHUNKS]);
}

// lambda pass (rx_sph_pbd.cpp:699-822) over the list: reads positions xa, writes xl = (x, lambda).
// Adaptive solve: once a check met the tolerance the remaining passes exit at once.
// One slot: C = rho / rho0 - 1 in C_out.
template <bool kCoop, class Walk>
__device__ inline bool lambda_slot(uint i, const float4* __restrict__ xa, float4* __restrict__ xl, float* __restrict__ dens,
                                   const PbfK& k, Walk&& walk, float& C_out) {
    float3 p = f3(xa[i]);
    float d = 0.0f, sd = 0.0f;
    float3 gs = f3(0, 0, 0);
    if (!walk(p, [&](uint j) { return xa[j]; },
              [&](uint, const float4& qj) {
                  float3 r = p - f3(qj);
                  float r2 = dot3(r, r);
                  if (r2 <= k.s.h2) {
                      float q2 = k.s.h2 - r2;
                      d += q2 * q2 * q2;
                      if (r2 > 0.0f) {
                          float ri = rinv_of(r2);
                          float q = k.s.h - r2 * ri;
                          float3 g = (k.gfac * q * q * ri) * r;
                          gs += g;
                          sd += dot3(g, g);
                      }
                  }
              }))
        return false;
    d = wsum<kCoop>(d);
    sd = wsum<kCoop>(sd);
    gs = wsum3<kCoop>(gs);
    d *= k.s.mass * k.s.wpoly6;
    sd += dot3(gs, gs);
    float C = d * k.s.inv_rho0 - 1.0f;
    float lam = -C / (sd + k.eps);
    if (writer<kCoop>()) {
        xl[i] = make_float4(p.x, p.y, p.z, lam);
        dens[i] = d;
    }
    C_out = C;
    return true;
}

__global__ __launch_bounds__(kTpb) void k_lambda(const float4* __restrict__ xa, float4* __restrict__ xl, float* __restrict__ dens,
                                                 float* __restrict__ stats, int it, PbfK k, const uint* __restrict__ list,
                                                 const uint4* __restrict__ cnt, const uint* __restrict__ bhash,
                                                 const uint* __restrict__ off, int adaptive, const float4* __restrict__ xb) {
    if (adaptive && it > 1 && converged(stats, it - 1, k)) return;  // solve already converged
    bool valid;
    uint i = slot_of(k, valid);
    float disp2 = stats[ST_DISP2];
    const NearMovers nm = stage_movers(i, valid, k, disp2, stats, xb, xa, bhash, false);
    // the movers the next correction pass reads (it appends new ones while it runs)
    if (blockIdx.x == 0 && threadIdx.x == 0) stats[ST_MOVERS_PUB] = stats[ST_MOVERS];
    float err = 0.0f;
    float C;
    if (valid && lambda_slot<false>(i, xa, xl, dens, k,
                                    [&](float3 p, auto&& fetch, auto&& body) {
                                        return for_neighbors<2>(i, 0, k, list, cnt, bhash, off, disp2, p, xb, nm, fetch, body);
                                    },
                                    C))
        err = fmaxf(C, 0.0f);
    block_max_atomic(err, &stats[ST_ERR_ITER + it]);
    block_sum_atomic(err, &stats[ST_MPOS_ITER + it]);
}

__global__ __launch_bounds__(kTpb) void k_lambda_deferred(const float4* __restrict__ xa, float4* __restrict__ xl,
                                                          float* __restrict__ dens, float* __restrict__ stats, int it, PbfK k,
                                                          const uint* __restrict__ bhash, const uint* __restrict__ off,
                                                          const float4* __restrict__ xb) {
    const uint nq = defer_count(stats), nmov = __float_as_uint(stats[ST_MOVERS]);
    const float disp2 = stats[ST_DISP2];
    float err = 0.0f, perr = 0.0f;
    for (uint q = wave_index(); q < nq; q += wave_count()) {
        const uint i = k.defer[q];
        const float di = sqrtf(len2_fixed(f3(xa[i]) - f3(xb[i])));
        float C;
        lambda_slot<true>(i, xa, xl, dens, k,
                          [&](float3, auto&& fetch, auto&& body) {
                              coop_neighbors(i, di, k, bhash, off, disp2, xb, nmov, fetch, body);
                              return true;
                          },
                          C);
        if (writer<true>()) {
            err = fmaxf(err, fmaxf(C, 0.0f));
            perr += fmaxf(C, 0.0f);
        }
    }
    block_max_atomic(err, &stats[ST_ERR_ITER + it]);
    block_sum_atomic(perr, &stats[ST_MPOS_ITER + it]);
    defer_done(stats);
}

// Position correction (rx_sph_pbd.cpp:925-1026) + apply + wall (rx_sph_pbd.cpp:436-445): reads xl = (x, lambda),
// writes xa = clamp(x + dp).  Tracks the displacement from the list build.  In learned mode the pass is
// skipped once the previous check met the tolerance.
// One slot: mv / mv0 = squared displacement from the build position after / before the correction.
template <bool kCoop, class Walk>
__device__ inline bool dp_slot(uint i, const float4* __restrict__ xl, float4* __restrict__ xa, float4* __restrict__ xbuild,
                               float* __restrict__ lamacc, float* __restrict__ stats, int it, const PbfK& k, Walk&& walk,
                               float& mv, float& mv0) {
    float4 p4 = xl[i];
    float3 p = f3(p4);
    float li = p4.w;
    float dt2 = k.dt * k.dt;
    float3 acc = f3(0, 0, 0);
    if (!walk(p, [&](uint j) { return xl[j]; },
              [&](uint, const float4& q4) {
                  float3 r = p - f3(q4);
                  float r2 = dot3(r, r);
                  if (r2 <= k.s.h2 && r2 > 0.0f) {
                      float ri = rinv_of(r2);
                      float sc = 0.0f;
                      if (k.s.ap) {
                          float q2 = k.s.h2 - r2;
                          float ww = k.s.wpoly6 * q2 * q2 * q2 * k.s.inv_ap_wq;
                          sc = -k.s.ap_k * (k.s.ap_n == 4.0f ? pow4(ww) : powf(ww, k.s.ap_n)) * dt2;
                      }
                      float q = k.s.h - r2 * ri;
                      acc += ((li + q4.w + sc) * k.gfac * q * q * ri) * r;
                  }
              }))
        return false;
    acc = wsum3<kCoop>(acc);
    if (it == 0 && k.first_dmax > 0.0f) {  // the warm start's correction, held to first_dmax
        const float l2 = dot3(acc, acc);
        if (l2 > k.first_dmax * k.first_dmax) acc = (k.first_dmax * rinv_of(l2)) * acc;
    }
    float3 xn = clamp_wall(p + acc, k.s);
    const float4 b4 = xbuild[i];
    mv = len2_fixed(xn - f3(b4));
    mv0 = len2_fixed(p - f3(b4));
    if (writer<kCoop>()) {
        xa[i] = make_float4(xn.x, xn.y, xn.z, 0.0f);
        lamacc[i] += li;  // the multiplier this step has applied so far
        if (mv > k.half_skin2 && b4.w == 0.0f) {  // a new mover: listed once until the next build
            xbuild[i].w = 1.0f;
            const uint t = atomicAdd((unsigned int*)&stats[ST_MOVERS], 1u);
            if (t < kMoverCap) k.movers[t] = make_float4(b4.x, b4.y, b4.z, __uint_as_float(i));
        }
    }
    return true;
}

__global__ __launch_bounds__(kTpb) void k_dp(const float4* __restrict__ xl, float4* __restrict__ xa, float4* __restrict__ xbuild,
                                             float* __restrict__ lamacc, float* __restrict__ stats, int it, PbfK k,
                                             const uint* __restrict__ list, const uint4* __restrict__ cnt,
                                             const uint* __restrict__ bhash, const uint* __restrict__ off, int adaptive) {
    if (adaptive && it > 0 && converged(stats, it, k)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        atomicAdd((unsigned int*)&stats[ST_ITERS], 1u);
        atomicAdd((unsigned int*)&stats[ST_RUN_ITERS], 1u);
    }
    bool valid;
    uint i = slot_of(k, valid);
    float disp2 = stats[ST_DISP2];
    const NearMovers nm = stage_movers(i, valid, k, disp2, stats, xbuild, xl, bhash, true);
    float mv = 0.0f, mv0 = 0.0f;
    float a, b;
    if (valid && dp_slot<false>(i, xl, xa, xbuild, lamacc, stats, it, k,
                                [&](float3 p, auto&& fetch, auto&& body) {
                                    return for_neighbors<1>(i, 0, k, list, cnt, bhash, off, disp2, p, xbuild, nm, fetch, body);
                                },
                                a, b)) {
        mv = a;
        mv0 = b;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {  // one atomic per wavefront and level, when a lane crossed
        const uint64_t c = __ballot(mv0 <= k.cross2[l] && mv > k.cross2[l]);
        if (c && __lane_id() == (uint)__builtin_ctzll(c)) atomicAdd((unsigned int*)&stats[ST_CROSS + l], (uint)__builtin_popcountll(c));
    }
    block_max_atomic(mv, &stats[ST_DISP2]);
}

__global__ __launch_bounds__(kTpb) void k_dp_deferred(const float4* __restrict__ xl, float4* __restrict__ xa,
                                                      float4* __restrict__ xbuild, float* __restrict__ lamacc,
                                                      float* __restrict__ stats, int it, PbfK k, const uint* __restrict__ bhash,
                                                      const uint* __restrict__ off) {
    // the movers as of the lambda pass, as the correction pass read them
    const uint nq = defer_count(stats), nmov = __float_as_uint(stats[ST_MOVERS_PUB]);
    const float disp2 = stats[ST_DISP2];
    float mvmax = 0.0f;
    for (uint q = wave_index(); q < nq; q += wave_count()) {
        const uint i = k.defer[q];
        const float di = sqrtf(len2_fixed(f3(xl[i]) - f3(xbuild[i])));
        float mv, mv0;
        dp_slot<true>(i, xl, xa, xbuild, lamacc, stats, it, k,
                      [&](float3, auto&& fetch, auto&& body) {
                          coop_neighbors(i, di, k, bhash, off, disp2, xbuild, nmov, fetch, body);
                          return true;
                      },
                      mv, mv0);
        if (writer<true>()) {
            mvmax = fmaxf(mvmax, mv);
            for (int l = 0; l < 4; ++l)
                if (mv0 <= k.cross2[l] && mv > k.cross2[l]) atomicAdd((unsigned int*)&stats[ST_CROSS + l], 1u);
        }
    }
    block_max_atomic(mvmax, &stats[ST_DISP2]);
    defer_done(stats);
}

// v = (x* - x_n)/dt (rx_sph_pbd.cpp:465-471), with rho in w for the XSPH pass.
__global__ void k_velocity(const float4* __restrict__ xa, const float4* __restrict__ x0, const float* __restrict__ dens,
                           float4* __restrict__ v, PbfK k) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.n) return;
    float3 vv = k.inv_dt * (f3(xa[i]) - f3(x0[i]));
    v[i] = make_float4(vv.x, vv.y, vv.z, dens[i]);
}

// XSPH (rx_pbdsph_kernel.cu:1090-1162 form): reads v = (v, rho), writes vout.  Held to at least six wavefronts
// per SIMD (it needs 68 registers, so seven run): the pass waits on neighbour gathers, and more wavefronts hide
// more of it.  Eight (64 registers) spill and run slower.
template <bool kCoop, class Walk>
__device__ inline bool xsph_slot(uint i, const float4* __restrict__ xa, const float4* __restrict__ v, float4* __restrict__ vout,
                                 const PbfK& k, Walk&& walk) {
    float3 p = f3(xa[i]);
    float3 vi = f3(v[i]);
    float3 acc = f3(0, 0, 0);
    if (!walk(p, [&](uint j) { return xa[j]; },
              [&](uint j, const float4& x4) {
                  float3 r = p - f3(x4);
                  float r2 = dot3(r, r);
                  if (r2 <= k.s.h2) {
                      const float4 vj = v[j];  // read for the neighbours within h only
                      float q = k.s.h2 - r2;
                      acc += (q * q * q / vj.w) * (f3(vj) - vi);
                  }
              }))
        return false;
    acc = wsum3<kCoop>(acc);
    float3 vn = vi + (k.xsph_c * k.s.mass * k.s.wpoly6) * acc;
    if (writer<kCoop>()) vout[i] = make_float4(vn.x, vn.y, vn.z, 0.0f);
    return true;
}

__global__ __launch_bounds__(kTpb) __attribute__((amdgpu_waves_per_eu(6))) void k_xsph(const float4* __restrict__ xa, const float4* __restrict__ v, float4* __restrict__ vout,
                                               float* __restrict__ stats, PbfK k, const uint* __restrict__ list,
                                               const uint4* __restrict__ cnt, const uint* __restrict__ bhash,
                                               const uint* __restrict__ off, const float4* __restrict__ xb) {
    bool valid;
    uint i = slot_of(k, valid);
    float disp2 = stats[ST_DISP2];
    const NearMovers nm = stage_movers(i, valid, k, disp2, stats, xb, xa, bhash, false);
    if (!valid) return;
    xsph_slot<false>(i, xa, v, vout, k, [&](float3 p, auto&& fetch, auto&& body) {
        return for_neighbors<1>(i, 0, k, list, cnt, bhash, off, disp2, p, xb, nm, fetch, body);
    });
}

__global__ __launch_bounds__(kTpb) void k_xsph_deferred(const float4* __restrict__ xa, const float4* __restrict__ v,
                                                        float4* __restrict__ vout, float* __restrict__ stats, PbfK k,
                                                        const uint* __restrict__ bhash, const uint* __restrict__ off,
                                                        const float4* __restrict__ xb) {
    const uint nq = defer_count(stats), nmov = __float_as_uint(stats[ST_MOVERS]);
    const float disp2 = stats[ST_DISP2];
    for (uint q = wave_index(); q < nq; q += wave_count()) {
        const uint i = k.defer[q];
        const float di = sqrtf(len2_fixed(f3(xa[i]) - f3(xb[i])));
        xsph_slot<true>(i, xa, v, vout, k, [&](float3, auto&& fetch, auto&& body) {
            coop_neighbors(i, di, k, bhash, off, disp2, xb, nmov, fetch, body);
            return true;
        });
    }
    defer_done(stats);
}

// ------------------------------------------------------------------------------------------ heat
// HeatTransfar::heatParticleAndParticle (HeatTransfar.cpp:429-473): over the neighbours within h at the
// final positions, dT_i = td sum_{j != i} dt (T_j - T_i) / max(rho_j, 0.1) W_spline(r_ij, h), with the
// cubic spline of HeatTransfar.cpp:200-236 (sigma = 10 / (7 pi h^2)).  therm = (T, heat, phase, changed).
// hp: (T_j, max(rho_j, 0.1)) per slot (k_heat_pack), read for the neighbours within h.
__global__ void k_heat_pack(const float4* __restrict__ therm, const float* __restrict__ dens, float2* __restrict__ hp, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hp[i] = make_float2(therm[i].x, fmaxf(dens[i], 0.1f));
}
template <bool kCoop, class Walk>
__device__ inline bool heat_slot(uint i, const float4* __restrict__ x, const float2* __restrict__ hp,
                                 const float4* __restrict__ therm, float* __restrict__ dtemp, const PbfK& k, float td, float hdt,
                                 Walk&& walk) {
    float3 p = f3(x[i]);
    float ti = therm[i].x;
    const float inv_h = k.inv_h, ws = 10.0f / (7.0f * 3.14159265358979f * k.s.h2);
    float acc = 0.0f;
    if (!walk(p, [&](uint j) { return x[j]; },
              [&](uint j, const float4& qj) {
                  float3 r = p - f3(qj);
                  float r2 = dot3(r, r);
                  if (r2 <= k.s.h2 && j != i) {
                      const float2 t = hp[j];  // one gather for the neighbour's temperature and density
                      float q = sqrtf(r2) * inv_h;
                      float w = q < 1.0f ? ws * (1.0f - 1.5f * q * q + 0.75f * q * q * q) : ws * 0.25f * (2.0f - q) * (2.0f - q) * (2.0f - q);
                      acc += hdt * (t.x - ti) / t.y * w;
                  }
              }))
        return false;
    acc = wsum<kCoop>(acc);
    if (writer<kCoop>()) dtemp[i] = acc * td;
    return true;
}

__global__ __launch_bounds__(kTpb) void k_heat_diffuse(const float4* __restrict__ x, const float2* __restrict__ hp,
                                                       const float4* __restrict__ therm, float* __restrict__ dtemp,
                                                       const float* __restrict__ stats, PbfK k, const uint* __restrict__ list,
                                                       const uint4* __restrict__ cnt, const uint* __restrict__ bhash,
                                                       const uint* __restrict__ off, const float4* __restrict__ xb, float td,
                                                       float hdt) {
    bool valid;
    uint i = slot_of(k, valid);
    float disp2 = stats[ST_DISP2];
    const NearMovers nm = stage_movers(i, valid, k, disp2, stats, xb, x, bhash, false);
    if (!valid) return;
    heat_slot<false>(i, x, hp, therm, dtemp, k, td, hdt, [&](float3 p, auto&& fetch, auto&& body) {
        return for_neighbors(i, 0, k, list, cnt, bhash, off, disp2, p, xb, nm, fetch, body);
    });
}

__global__ __launch_bounds__(kTpb) void k_heat_deferred(const float4* __restrict__ x, const float2* __restrict__ hp,
                                                        const float4* __restrict__ therm, float* __restrict__ dtemp,
                                                        float* __restrict__ stats, PbfK k, const uint* __restrict__ bhash,
                                                        const uint* __restrict__ off, const float4* __restrict__ xb, float td,
                                                        float hdt) {
    const uint nq = defer_count(stats), nmov = __float_as_uint(stats[ST_MOVERS]);
    const float disp2 = stats[ST_DISP2];
    for (uint q = wave_index(); q < nq; q += wave_count()) {
        const uint i = k.defer[q];
        const float di = sqrtf(len2_fixed(f3(x[i]) - f3(xb[i])));
        heat_slot<true>(i, x, hp, therm, dtemp, k, td, hdt, [&](float3, auto&& fetch, auto&& body) {
            coop_neighbors(i, di, k, bhash, off, disp2, xb, nmov, fetch, body);
            return true;
        });
    }
    defer_done(stats);
}

// HeatTransfar::calcTempAndHeat (HeatTransfar.cpp:280-350): latent / sensible heat and the phase state
// (-2 ice, -1 melting, 1 freezing, 2 water; 250 is the melting point of the reference's temperature scale).
__global__ void k_heat_phase(float4* __restrict__ therm, const float* __restrict__ dtemp, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float latent = 165.0f;
    float4 th = therm[i];
    float t = th.x, q = th.y, dT = dtemp[i];
    int ph = (int)th.z;
    float chg = th.w;
    if (ph == -2 && t > 250.0f) {
        ph = -1; t = 250.0f; q = 0.0f;
    } else if (ph == 2 && t < 250.0f) {
        ph = 1; t = 250.0f; q = latent;
    }
    if (ph == -1 || ph == 1) {
        float c = 2.1f + (2.1f * q / latent);
        q += dT * c * 1.0f;
        if (ph == -1 && q < -latent / 300.0f) {
            ph = -2; t = 249.0f; q = 0.0f;
        } else if (ph == 1 && q > latent) {
            ph = 2; t = 251.0f; q = 0.0f;
        }
        if (ph == -1 && q > latent) {
            ph = 2; t = 251.0f; q = 0.0f; chg = 1.0f;
        } else if (ph == 1 && q < 0.0f) {
            ph = -2; t = 249.0f; q = 0.0f; chg = 1.0f;
        }
    } else {
        float c = (t > 250.0f) ? 4.2f : 2.1f;
        t = t + q / (c * 1.0f);
        q = 0.0f;
        t += dT;
        t = fminf(fmaxf(t, 0.0f), 500.0f);
    }
    therm[i] = make_float4(t, q, (float)ph, chg);
}

__global__ void k_gather_by_id(float4* __restrict__ dst, const float4* __restrict__ src, const uint* __restrict__ id, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[id[i]];
}

__global__ void k_scatter_out(const float4* __restrict__ src, const uint* __restrict__ id, float4* __restrict__ dst, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[id[i]] = src[i];
}
__global__ void k_scatter_out1(const float* __restrict__ src, const uint* __restrict__ id, float* __restrict__ dst, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[id[i]] = src[i];
}
__global__ void k_scatter_feat(const float4* __restrict__ src, const uint* __restrict__ id, float4* __restrict__ dst, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint o = id[i];
    for (int t = 0; t < 4; ++t) dst[4 * (size_t)o + t] = src[4 * (size_t)i + t];
}
__global__ void k_iota(uint* __restrict__ a, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

}  // namespace

// ============================================================================================ class
rxPBDSPH_GPU::rxPBDSPH_GPU() {}
rxPBDSPH_GPU::~rxPBDSPH_GPU() { Finalize(); }

void rxPBDSPH_GPU::Finalize() {
    if (m_stream) RX_HIP(hipStreamSynchronize(m_stream));
    drop_graphs();
    for (void** p : {(void**)&m_x, (void**)&m_xs, (void**)&m_xl, (void**)&m_xb, (void**)&m_v, (void**)&m_t, (void**)&m_dens,
                     (void**)&m_lamacc, (void**)&m_id, (void**)&m_idt, (void**)&m_list, (void**)&m_cnt, (void**)&m_stats,
                     (void**)&m_ftmp, (void**)&m_feat, (void**)&m_ftmp16, (void**)&m_therm, (void**)&m_therm_t, (void**)&m_mlp_mem,
                     (void**)&m_movers, (void**)&m_defer})
        if (*p) {
            RX_HIP(hipFree(*p));
            *p = nullptr;
        }
    m_mlp_ready = false;
    m_grid.release();
    if (m_hrec) {
        RX_HIP(hipHostFree(m_hrec));
        m_hrec = nullptr;
    }
    for (hipEvent_t& e : m_rec_ev)
        if (e) {
            RX_HIP(hipEventDestroy(e));
            e = nullptr;
        }
    m_rec_head = m_rec_tail = 0;
    for (hipEvent_t e : m_event_pool) RX_HIP(hipEventDestroy(e));
    m_event_pool.clear();
    m_marks.clear();
    m_event_next = 0;
    if (m_stream) {
        RX_HIP(hipStreamDestroy(m_stream));
        m_stream = nullptr;
    }
    m_n = 0;
    m_have_list = false;
}

// Bytes of the skin list for n slots of capacity maxn: ceil(maxn / 8) uint4 chunks of 8 16-bit entries per slot.
static size_t list_bytes(size_t n, int maxn) { return sizeof(uint4) * n * (size_t)((maxn + 7) / 8); }

int rxPBDSPH_GPU::Initialize(const rxSphEnv& env) {
    Finalize();
    if (rxSetupSimParams(&env, &m_params) != 0) return 1;
    m_env = env;
    m_skin = m_skin_base = 0.15f * env.h;
    m_level = m_level_target = m_calm = 0;
    m_maxn = m_maxn_target = 96;
    m_maxn_cap = 512;
    m_grid.setup(m_params, env.h + m_skin);
    lattice_refs();
    if (!RX_HIP(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking))) return 1;
    if (!RX_HIP(hipHostMalloc((void**)&m_hrec, sizeof(float) * kRecLen * kRecs, hipHostMallocDefault))) return 1;
    for (hipEvent_t& e : m_rec_ev)
        if (!RX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming))) return 1;
    return 0;
}

// Applies the step records (oldest first) and updates the targets.  Level want_d is the lowest whose skin is at
// least 2.2 times the step's largest displacement D: then d_i + D <= 2D stays within the skin and no particle
// walks the grid.  Level want_m(a) is the lowest at which at most a particles crossed half its skin: those few
// walk the grid as movers and the rest keep their lists, at a cost that stays small while they are sparse
// (allowance A = min(kMoverCap / 2, n / 256)).  A grid walk costs 10 to 20 list walks while each level up costs
// 1.4 to 1.9 times the list work, so the level rises only when more than 3% of the step's neighbour passes
// walked the grid for want of a wide enough skin (the walks of overflowed lists, one per pass, do not count: a
// wider skin only lengthens the lists): at once to min(want_d, want_m(A)), and at least one level (up to
// want_d) since the walks show the movers were not few enough.  It falls to min(want_d, want_m(A / 4)) (the
// lower allowance keeps it from falling back into the walks) after 2 records in a row want a lower level.  The
// list capacity rises by 64 when more than n/2000 lists overflowed.
void rxPBDSPH_GPU::poll_records() {
    // Before step s the records of steps up to s - 2 are applied, waiting for them if need be, and never the record
    // of step s - 1: the sizes of every step are then a function of the run alone, not of how far the host ran
    // ahead (the device still has step s - 1 queued while the host issues step s).
    while (m_rec_tail + 1 < m_rec_head) {
        const int slot = (int)(m_rec_tail % kRecs);
        if (!RX_HIP(hipEventSynchronize(m_rec_ev[slot]))) break;
        const float* r = m_hrec + kRecLen * slot;  // stats slots ST_DISP2 .. ST_CROSS + 3
        auto u = [&](int st) { return __builtin_bit_cast(uint, r[st - ST_DISP2]); };
        const float D = sqrtf(fmaxf(r[0], 0.0f));
        const uint over = u(ST_OVERFLOW), fb = u(ST_FALLBACK);
        ++m_rec_tail;
        if (u(ST_BAD)) {
            // a pass of that step walked less than the reach it needed: start the next step from a fresh grid and
            // let Update report it
            ++m_bad_steps;
            m_have_list = false;
        }
        if (!m_adapt) continue;
        const float need = 2.2f * D;
        int want_d = 0;
        while (want_d < 3 && m_skin_base * (float)(1 + want_d) < need) ++want_d;
        auto want_m = [&](uint allowance) {
            int l = 0;
            while (l < 3 && u(ST_CROSS + l) > allowance) ++l;
            return l;
        };
        const uint allowance = std::min(kMoverCap / 2, m_n / 256);
        const float passes = (float)(2 * std::min(m_env.max_iter, 14) + 2);
        const int want_up = std::max(std::min(want_d, want_m(allowance)), std::min(want_d, m_level_target + 1));
        const int want = std::min(want_d, want_m(allowance / 4));
        const float fb_skin = fmaxf((float)fb - (float)over * passes, 0.0f);  // walks a wider skin would spare
        if (fb_skin > 0.03f * (float)m_n * passes) {
            if (want_up > m_level_target) m_level_target = want_up;
            m_calm = 0;
        } else if (want < m_level_target) {
            m_calm_want = m_calm ? std::max(m_calm_want, want) : want;
            if (++m_calm >= 2) {
                m_level_target = m_calm_want;
                m_calm = 0;
            }
        } else {
            m_calm = 0;
        }
        if (over > m_n / 2000 && m_maxn_target < m_maxn_cap) m_maxn_target = std::min(m_maxn_cap, m_maxn_target + 64);
    }
}

// Copies the step's slots ST_DISP2 .. ST_CROSS + 3 behind its work, with an event (at most two records wait to be
// applied; the copy is skipped only if the ring were full).
void rxPBDSPH_GPU::push_record() {
    if (!m_hrec || m_rec_head - m_rec_tail >= kRecs) return;
    const int slot = (int)(m_rec_head % kRecs);
    if (RX_HIP(hipMemcpyAsync(m_hrec + kRecLen * slot, m_stats + ST_DISP2, kRecLen * sizeof(float), hipMemcpyDeviceToHost, m_stream)) &&
        RX_HIP(hipEventRecord(m_rec_ev[slot], m_stream)))
        ++m_rec_head;
}

// Applies the targets between the prediction (which read the previous list) and the rebuild.  A new capacity
// reallocates the list, so the passes still reading it are waited for first.  A new skin changes the cell of
// the grid to h + skin; the cells only grow relative to level 0, so the reserved cell arrays suffice.
bool rxPBDSPH_GPU::apply_targets(hipStream_t s) {
    bool changed = false;
    if (m_level_target > m_level) {
        // a wider skin holds more neighbours per list: the capacity grows with the volume of the cutoff at once
        const float r = (m_env.h + m_skin_base * (float)(1 + m_level_target)) / (m_env.h + m_skin_base * (float)(1 + m_level));
        const int grown = ((int)ceilf((float)m_maxn * r * r * r) + 31) & ~31;
        m_maxn_target = std::min(m_maxn_cap, std::max(m_maxn_target, grown));
    }
    if (m_maxn_target != m_maxn) {
        // the larger list is allocated before the old one is released; when the device has no room for it the
        // capacity stays (overflowed lists keep walking the grid) and no larger one is asked for again
        uint* grown = nullptr;
        if (hipMalloc((void**)&grown, list_bytes(m_n, m_maxn_target)) == hipSuccess) {
            drop_graphs();
            if (!RX_HIP(hipStreamSynchronize(s))) {
                RX_HIP(hipFree(grown));
                return false;
            }
            RX_HIP(hipFree(m_list));
            m_list = grown;
            m_maxn = m_maxn_target;
            changed = true;
        } else {
            (void)hipGetLastError();
            m_maxn_target = m_maxn;
            m_maxn_cap = m_maxn;
        }
    }
    if (m_level_target != m_level) {
        m_level = m_level_target;
        m_skin = m_skin_base * (float)(1 + m_level);
        m_grid.setup(m_params, m_env.h + m_skin);
        changed = true;
    }
    return changed;
}


bool rxPBDSPH_GPU::alloc(uint n) {
    size_t v4 = sizeof(float4) * n, f = sizeof(float) * n, u = sizeof(uint) * n;
    bool ok = RX_HIP(hipMalloc(&m_x, v4)) && RX_HIP(hipMalloc(&m_xs, v4)) && RX_HIP(hipMalloc(&m_xl, v4)) &&
              RX_HIP(hipMalloc(&m_xb, v4)) && RX_HIP(hipMalloc(&m_v, v4)) && RX_HIP(hipMalloc(&m_t, v4)) &&
              RX_HIP(hipMalloc(&m_dens, f)) && RX_HIP(hipMalloc(&m_lamacc, f)) && RX_HIP(hipMalloc(&m_id, u)) &&
              RX_HIP(hipMalloc(&m_idt, u)) && RX_HIP(hipMalloc(&m_list, list_bytes(n, m_maxn))) && RX_HIP(hipMalloc(&m_cnt, v4)) &&
              RX_HIP(hipMalloc(&m_stats, sizeof(float) * ST_N)) && RX_HIP(hipMalloc(&m_ftmp, f)) &&
              RX_HIP(hipMalloc(&m_feat, v4 * 4)) &&
              RX_HIP(hipMalloc(&m_ftmp16, v4 * 4)) && RX_HIP(hipMalloc(&m_movers, sizeof(float4) * kMoverCap)) &&
              RX_HIP(hipMalloc(&m_defer, sizeof(uint) * kDeferCap));
    m_grid.n = n;
    m_grid.setup(m_params, m_env.h);  // the cell-h grid of the first prediction has the most cells
    ok = ok && m_grid.reserve(n);
    m_grid.setup(m_params, m_env.h + m_skin);
    return ok;
}

int rxPBDSPH_GPU::SetParticles(const float* pos4, const float* vel4, int n) {
    if (n <= 0 || !pos4) {
        rx_set_error(1, "rxPBDSPH_GPU::SetParticles: bad arguments");
        return 1;
    }
    if (n >= (1 << 30)) {  // the list build packs a plane number above 30-bit slot numbers
        rx_set_error(1, "rxPBDSPH_GPU::SetParticles: at most 2^30 - 1 particles");
        return 1;
    }
    drop_graphs();
    m_rec_tail = m_rec_head;  // records of the previous particles no longer apply
    m_bad_steps = m_bad_reported = 0;
    m_level_target = 0;
    m_calm = 0;
    if ((uint)n != m_n) {
        for (void** p : {(void**)&m_x, (void**)&m_xs, (void**)&m_xl, (void**)&m_xb, (void**)&m_v, (void**)&m_t, (void**)&m_dens,
                         (void**)&m_lamacc, (void**)&m_id, (void**)&m_idt, (void**)&m_list, (void**)&m_cnt, (void**)&m_stats,
                         (void**)&m_ftmp, (void**)&m_feat, (void**)&m_ftmp16, (void**)&m_therm, (void**)&m_therm_t, (void**)&m_movers,
                         (void**)&m_defer})
            if (*p) {
                RX_HIP(hipFree(*p));
                *p = nullptr;
            }
        m_n = (uint)n;
        if (!alloc(m_n)) {
            m_n = 0;  // Update refuses until a SetParticles succeeds
            return 1;
        }
    }
    hipStream_t s = m_stream;
    RX_HIP(hipMemcpyAsync(m_x, pos4, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    if (vel4) RX_HIP(hipMemcpyAsync(m_v, vel4, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    else rx_fill<float4>(m_v, make_float4(0, 0, 0, 0), n, s);
    k_iota<<<rx_blocks(n, 256), 256, 0, s>>>(m_id, n);
    rx_fill<float>(m_lamacc, 0.0f, n, s);
    rx_fill<float>(m_stats, 0.0f, ST_N, s);
    m_have_list = false;
    m_step = 0;
    return RX_HIP(hipStreamSynchronize(s)) ? 0 : 1;
}

// Lambda denominator and neighbour count of an interior particle of the rest lattice (spacing env.spacing):
// the scales of the learned warm start's features.
void rxPBDSPH_GPU::lattice_refs() {
    double h = m_env.h, l = m_env.spacing > 0.0f ? m_env.spacing : 0.5 * h;
    double gw = -45.0 / (3.14159265358979323846 * pow(h, 6.0)), r0 = m_params.Density;
    const double gm = m_grad_mass ? (double)m_params.Mass : 1.0;
    int nmax = (int)ceil(h / l) + 1;
    double sd = 0.0;
    int cnt = 0;
    for (int x = -nmax; x <= nmax; ++x)
        for (int y = -nmax; y <= nmax; ++y)
            for (int z = -nmax; z <= nmax; ++z) {
                double r = sqrt((double)(x * x + y * y + z * z)) * l;
                if (r > h) continue;
                ++cnt;
                if (r > 0.0) {
                    double g = gw * (h - r) * (h - r) / r0 * gm;  // |grad W| / rho0 (times m in the mass mode)
                    sd += g * g;
                }
            }
    m_sd_ref = (float)sd;
    m_nref = (float)cnt;
}

PbfK rxPBDSPH_GPU::consts(float dt) const {
    PbfK k;
    memset(&k, 0, sizeof(k));  // padding included: graph keys compare the bytes
    rxSimParams p = m_params;
    p.Dt = dt;
    k.s = rx_make_consts(p);
    k.g = m_grid.k;
    k.n = m_n;
    float rc = m_env.h + m_skin;
    k.rc2 = rc * rc;
    k.half_skin2 = 0.25f * m_skin * m_skin;
    k.skin = m_skin;
    float wide = 0.5f * ((float)kMaxStencil * m_grid.k.cell - m_env.h);  // largest fallback reach minus h, per particle
    k.wide2 = wide * wide;
    k.fallback = (uint*)(m_stats + ST_FALLBACK);
    k.bad = (uint*)(m_stats + ST_BAD);
    k.movers = m_movers;
    k.defer = m_defer;
    k.defer_n = (uint*)(m_stats + ST_DEFER);
    k.defer_on = m_defer_on;
    for (int l = 0; l < 4; ++l) {
        const float t = 0.5f * m_skin_base * (float)(1 + l);
        k.cross2[l] = t * t;
    }
    k.maxn = m_maxn;
    k.eps = m_env.epsilon;
    k.gfac = k.s.gwspiky * k.s.inv_rho0 * (m_grad_mass ? k.s.mass : 1.0f);
    k.dt = dt;
    k.inv_dt = 1.0f / dt;
    k.xsph_c = m_env.xsph_c;
    k.tol = m_tol;
    k.judge = m_judge;
    k.lam0_scale = m_mode == RX_PBF_MODE_SCALED ? m_warm_scale : 1.0f;
    k.warm_kappa = m_mode == RX_PBF_MODE_LEARNED ? m_warm_kappa : 0.0f;
    k.first_dmax = m_first_dmax * m_env.h;
    k.sd_ref = m_sd_ref;
    k.inv_nref = 1.0f / m_nref;
    k.inv_gref = 1.0f / sqrtf(m_sd_ref);
    k.inv_h = 1.0f / m_env.h;
    k.inv_spacing = 1.0f / (m_env.spacing > 0.0f ? m_env.spacing : 0.5f * m_env.h);
    return k;
}

// Gathers every per-particle state array into the order of the last grid build.
void rxPBDSPH_GPU::reorder(float4*& a, hipStream_t s) {
    fg_gather<float4>(m_t, a, m_grid.perm, m_n, s);
    std::swap(a, m_t);
}

rxPBDSPH_GPU::StepPtrs rxPBDSPH_GPU::save_ptrs() const {
    return StepPtrs{m_x, m_xs, m_xl, m_xb, m_v, m_t, m_therm, m_therm_t, m_dens, m_lamacc, m_ftmp, m_id, m_idt, m_grid.perm, m_grid.slot};
}
void rxPBDSPH_GPU::load_ptrs(const StepPtrs& p) {
    m_x = p.x, m_xs = p.xs, m_xl = p.xl, m_xb = p.xb, m_v = p.v, m_t = p.t, m_therm = p.therm, m_therm_t = p.therm_t;
    m_dens = p.dens, m_lamacc = p.lamacc, m_ftmp = p.ftmp, m_id = p.id, m_idt = p.idt, m_grid.perm = p.perm, m_grid.slot = p.slot;
}
// Everything the launches of one step take from the host: buffers, the kernel constants, the network view and
// the choices the step makes on the host.
std::vector<unsigned char> rxPBDSPH_GPU::graph_key(float dt) const {
    std::vector<unsigned char> key;
    auto put = [&](const void* p, size_t n) { key.insert(key.end(), (const unsigned char*)p, (const unsigned char*)p + n); };
    StepPtrs ptrs = save_ptrs();
    put(&ptrs, sizeof(ptrs));
    PbfK k = consts(dt);
    put(&k, sizeof(k));
    put(&m_mlp, sizeof(m_mlp));
    const int flags[] = {m_mode, std::min(m_env.max_iter, 14), m_adaptive ? 1 : 0, m_list_tile, m_record_features ? 1 : 0,
                         m_step > 0 ? 1 : 0, m_have_list ? 1 : 0};
    put(flags, sizeof(flags));
    const float vals[] = {m_warm_scale, m_heat_td, m_heat_dt > 0.0f ? m_heat_dt : dt};
    put(vals, sizeof(vals));
    const void* bufs[] = {m_list, m_cnt, m_stats, m_feat, m_grid.hash, m_grid.off};
    put(bufs, sizeof(bufs));
    return key;
}
void rxPBDSPH_GPU::drop_graphs() {
    if (m_graphs.empty()) return;
    RX_HIP(hipStreamSynchronize(m_stream));  // no launch of these graphs is still pending
    for (StepGraph& g : m_graphs) RX_HIP(hipGraphExecDestroy(g.exec));
    m_graphs.clear();
}

int rxPBDSPH_GPU::Update(float dt, int step) {
    if (!m_n) {
        rx_set_error(1, "rxPBDSPH_GPU::Update: no particles");
        return 1;
    }
    if (m_mode == RX_PBF_MODE_LEARNED && !m_mlp_ready) {
        rx_set_error(1, "rxPBDSPH_GPU::Update: learned mode needs network weights (rxPBDSPH_GPU_SetNetwork)");
        return 1;
    }
    int rc = 0;
    poll_records();
    if (m_bad_steps > m_bad_reported) {
        m_bad_reported = m_bad_steps;
        rx_set_error(3, "rxPBDSPH_GPU::Update: an earlier step moved particles beyond the %d-cell reach of the "
                        "neighbour walk, so that step missed pairs; the neighbour list is rebuilt from this step on",
                     kMaxStencil);
        return 3;
    }
    if (!m_use_graph || m_profile || !m_have_list) {
        rc = enqueue_step(dt);  // the first step (no list yet) and profiled steps run eagerly
        if (rc == 0) push_record();
    } else {
        m_level_target = m_level;  // replayed and captured steps keep the sizes they were recorded with
        m_maxn_target = m_maxn;
        std::vector<unsigned char> key = graph_key(dt);
        StepGraph* hit = nullptr;
        for (StepGraph& g : m_graphs)
            if (g.key == key) hit = &g;
        if (hit) {
            if (!RX_HIP(hipGraphLaunch(hit->exec, m_stream))) return 1;
            load_ptrs(hit->after);
            m_have_list = true;
            ++m_step;
            push_record();
        } else {
            // capture this step's launches, then run it from the graph
            if (m_graphs.size() >= 12) drop_graphs();
            const StepPtrs before = save_ptrs();
            const int step0 = m_step;
            hipGraph_t g = nullptr;
            if (!RX_HIP(hipStreamBeginCapture(m_stream, hipStreamCaptureModeThreadLocal))) return 1;
            rc = enqueue_step(dt);
            const bool ended = RX_HIP(hipStreamEndCapture(m_stream, &g));
            hipGraphExec_t exec = nullptr;
            const bool ok = rc == 0 && ended && g && RX_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
            if (g) RX_HIP(hipGraphDestroy(g));
            if (!ok) {  // undo the host side of the captured step and run it eagerly
                load_ptrs(before);
                m_step = step0;
                m_use_graph = false;
                rc = enqueue_step(dt);
            } else {
                m_graphs.push_back(StepGraph{key, save_ptrs(), exec});
                if (!RX_HIP(hipGraphLaunch(exec, m_stream))) return 1;
            }
            push_record();
        }
    }
    if (rc) return rc;
    if (!RX_HIP(hipGetLastError())) return 1;
    if (m_sync_each_step && !RX_HIP(hipStreamSynchronize(m_stream))) return 1;
    return 0;
}

// The launches of one step on the solver's stream (eager, or recorded into a graph by Update).
int rxPBDSPH_GPU::enqueue_step(float dt) {
    hipStream_t s = m_stream;
    PbfK k = consts(dt);
    uint n = m_n;
    unsigned nb = rx_blocks(n, kTpb);
    const int K = std::min(m_env.max_iter, 14);
    const int adaptive = m_adaptive ? 1 : 0;
    // per-step statistics (ST_DISP2 and the movers carry over: the prediction walks the previous list)
    k_zero_slots<<<1, 64, 0, s>>>(m_stats, ((1ull << ST_RUN) - 1ull) & ~(1ull << ST_DISP2) & ~(1ull << ST_MOVERS));

    // ---- predict at x_n
    int mode = 0;
    if (!m_have_list) {
        // no list yet: a grid of cell h on x_n (the smallest cell a 3x3x3 stencil allows: 27 cells of about
        // 8 particles instead of about 12 at the list's cell h + skin), gather, predict with its stencil
        m_grid.setup(m_params, m_env.h);
        k = consts(dt);
        m_grid.build(m_x, s);
        reorder(m_x, s);
        reorder(m_v, s);
        fg_gather<uint>(m_idt, m_id, m_grid.perm, n, s);
        std::swap(m_id, m_idt);
        fg_gather<float>(m_ftmp, m_lamacc, m_grid.perm, n, s);
        std::swap(m_lamacc, m_ftmp);
        if (m_therm) {
            fg_gather<float4>(m_therm_t, m_therm, m_grid.perm, n, s);
            std::swap(m_therm, m_therm_t);
        }
        mode = 1;
    }
    int mk = begin_mark(RX_PROF_PREDICT);
    k_predict<<<nb, kTpb, 0, s>>>(m_x, m_v, m_xs, m_dens, m_stats, mode, k, m_list, m_cnt, m_grid.hash, m_grid.off, m_xb);
    end_mark(mk);
    if (mode == 0 && m_defer_on) {
        mk = begin_mark(RX_PROF_DEFERRED);
        k_predict_deferred<<<kDeferBlocks, kTpb, 0, s>>>(m_x, m_v, m_xs, m_dens, m_stats, k, m_grid.hash, m_grid.off, m_xb);
        end_mark(mk);
    }
    if (m_step > 0) k_fold_run<<<1, 64, 0, s>>>(m_stats, k);
    if (mode == 1) {  // back to the list's grid
        m_grid.setup(m_params, m_env.h + m_skin);
        k = consts(dt);
    }
    if (apply_targets(s)) k = consts(dt);  // list sizes for this step's build (the prediction used the old list)

    // ---- rebuild on x*; the step-start density follows the particles
    k_zero_slots<<<1, 64, 0, s>>>(m_stats, (1ull << ST_DISP2) | (1ull << ST_MOVERS) | (1ull << ST_MOVERS_PUB) | (1ull << ST_MOVERS_SORTED));
    mk = begin_mark(RX_PROF_GRID);
    {
        // the sort's last pass gathers the whole state into the new order (xl and xb are dead until the list
        // pass rewrites them, so they take the gathered positions and velocities) and clears the multiplier sums
        RadixGather gs;
        gs.n4 = m_therm ? 4 : 3;
        gs.src4[0] = m_xs, gs.dst4[0] = m_t;
        gs.src4[1] = m_x, gs.dst4[1] = m_xl;
        gs.src4[2] = m_v, gs.dst4[2] = m_xb;
        gs.src4[3] = m_therm, gs.dst4[3] = m_therm_t;
        gs.srcu = m_id, gs.dstu = m_idt;
        gs.zero = m_lamacc;
        m_grid.build(m_xs, s, &gs);
    }
    std::swap(m_xs, m_t);
    std::swap(m_x, m_xl);
    std::swap(m_v, m_xb);
    std::swap(m_id, m_idt);
    if (m_therm) std::swap(m_therm, m_therm_t);
    end_mark(mk);

    // ---- list + lambda_0 (the network's multiplier in learned mode), then up to K corrections
    const bool learned = m_mode == RX_PBF_MODE_LEARNED;
    // the default list pass runs the network itself; the LDS-tiled variants write the features for a separate
    // inference launch
    const bool fused_net = learned && m_list_tile == 0;
    const bool feats = (learned && !fused_net) || m_record_features;
    mk = begin_mark(learned || m_record_features ? RX_PROF_FEATURES : RX_PROF_LIST_LAMBDA);
    {
        float* f = feats ? m_feat : nullptr;
        const FGridK& gk = m_grid.k;
        switch (m_list_tile) {
        case 1:
            k_list_lambda<16, 3072, 256><<<list_segments<16>(gk), 256, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k,
                                                                             m_grid.hash, m_grid.off, f, m_v);
            break;
        case 2:
            k_list_lambda<16, 2048, 256><<<list_segments<16>(gk), 256, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k,
                                                                             m_grid.hash, m_grid.off, f, m_v);
            break;
        case 3:
            k_list_lambda<8, 1536, 128><<<list_segments<8>(gk), 128, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k,
                                                                           m_grid.hash, m_grid.off, f, m_v);
            break;
        case 4:  // segments without staging: every segment reads its candidates from global memory
            k_list_lambda<16, 0, 256><<<list_segments<16>(gk), 256, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k,
                                                                          m_grid.hash, m_grid.off, f, m_v);
            break;
        default:
            if (fused_net)
                k_list_lambda_slots<true><<<nb, kTpb, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k, m_grid.hash, m_grid.off, f,
                                                              m_v, m_mlp, m_warm_scale / k.sd_ref);
            else
                k_list_lambda_slots<false><<<nb, kTpb, 0, s>>>(m_xs, m_xl, m_xb, m_dens, m_list, m_cnt, m_stats, k, m_grid.hash, m_grid.off,
                                                               f, m_v, MlpDev{}, 0.0f);
        }
    }
    end_mark(mk);
    if (learned && !fused_net) {
        mk = begin_mark(RX_PROF_MLP);
        rx_mlp_infer_to_w(m_feat, m_xl, n, m_mlp, m_warm_scale / k.sd_ref, k.warm_kappa, s);
        end_mark(mk);
    }
    for (int it = 0; it < K; ++it) {
        if (it > 0) {
            mk = begin_mark(RX_PROF_LAMBDA);
            k_lambda<<<nb, kTpb, 0, s>>>(m_xs, m_xl, m_dens, m_stats, it, k, m_list, m_cnt, m_grid.hash, m_grid.off, adaptive, m_xb);
            end_mark(mk);
            if (m_defer_on) {
                mk = begin_mark(RX_PROF_DEFERRED);
                k_lambda_deferred<<<kDeferBlocks, kTpb, 0, s>>>(m_xs, m_xl, m_dens, m_stats, it, k, m_grid.hash, m_grid.off, m_xb);
                end_mark(mk);
            }
        }
        mk = begin_mark(RX_PROF_DP);
        k_dp<<<nb, kTpb, 0, s>>>(m_xl, m_xs, m_xb, m_lamacc, m_stats, it, k, m_list, m_cnt, m_grid.hash, m_grid.off, adaptive);
        end_mark(mk);
        mk = begin_mark(RX_PROF_DEFERRED);
        if (m_defer_on) k_dp_deferred<<<kDeferBlocks, kTpb, 0, s>>>(m_xl, m_xs, m_xb, m_lamacc, m_stats, it, k, m_grid.hash, m_grid.off);
        k_sort_movers<<<1, kSortThreads, 0, s>>>(m_movers, m_stats);
        end_mark(mk);
    }
    // (the density of the final positions is judged by the next step's prediction pass, slot ST_ERR_START)
    // ---- velocity, XSPH
    mk = begin_mark(RX_PROF_VELOCITY);
    k_velocity<<<rx_blocks(n, 256), 256, 0, s>>>(m_xs, m_x, m_dens, m_t, k);
    end_mark(mk);
    if (m_env.xsph_c != 0.0f) {
        mk = begin_mark(RX_PROF_XSPH);
        k_xsph<<<nb, kTpb, 0, s>>>(m_xs, m_t, m_v, m_stats, k, m_list, m_cnt, m_grid.hash, m_grid.off, m_xb);
        end_mark(mk);
        if (m_defer_on) {
            mk = begin_mark(RX_PROF_DEFERRED);
            k_xsph_deferred<<<kDeferBlocks, kTpb, 0, s>>>(m_xs, m_t, m_v, m_stats, k, m_grid.hash, m_grid.off, m_xb);
            end_mark(mk);
        }
    } else {
        std::swap(m_v, m_t);
    }
    std::swap(m_x, m_xs);
    if (m_therm) {  // heat exchange at the final positions
        mk = begin_mark(RX_PROF_HEAT);
        k_heat_pack<<<rx_blocks(n, 256), 256, 0, s>>>(m_therm, m_dens, (float2*)m_t, n);  // m_t is free once the velocity is final
        k_heat_diffuse<<<nb, kTpb, 0, s>>>(m_x, (const float2*)m_t, m_therm, m_ftmp, m_stats, k, m_list, m_cnt, m_grid.hash, m_grid.off, m_xb,
                                           m_heat_td, m_heat_dt > 0.0f ? m_heat_dt : dt);
        if (m_defer_on)
            k_heat_deferred<<<kDeferBlocks, kTpb, 0, s>>>(m_x, (const float2*)m_t, m_therm, m_ftmp, m_stats, k, m_grid.hash, m_grid.off, m_xb,
                                                          m_heat_td, m_heat_dt > 0.0f ? m_heat_dt : dt);
        k_heat_phase<<<rx_blocks(n, 256), 256, 0, s>>>(m_therm, m_ftmp, n);
        end_mark(mk);
    }
    m_have_list = true;
    ++m_step;
    return 0;
}

int rxPBDSPH_GPU::GetArray(int type, float* host, int count) {
    if (!host || count < (int)m_n) {
        rx_set_error(1, "rxPBDSPH_GPU::GetArray: buffer too small");
        return 1;
    }
    hipStream_t s = m_stream;
    uint n = m_n;
    float4* tmp = m_t;
    switch (type) {
    case RX_ARRAY_POSITION:
        k_scatter_out<<<rx_blocks(n, 256), 256, 0, s>>>(m_x, m_id, tmp, n);
        RX_HIP(hipMemcpyAsync(host, tmp, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        break;
    case RX_ARRAY_VELOCITY:
        k_scatter_out<<<rx_blocks(n, 256), 256, 0, s>>>(m_v, m_id, tmp, n);
        RX_HIP(hipMemcpyAsync(host, tmp, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        break;
    case RX_ARRAY_DENSITY:
        k_scatter_out1<<<rx_blocks(n, 256), 256, 0, s>>>(m_dens, m_id, (float*)tmp, n);
        RX_HIP(hipMemcpyAsync(host, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        break;
    case RX_ARRAY_LAMBDA_SUM:
        k_scatter_out1<<<rx_blocks(n, 256), 256, 0, s>>>(m_lamacc, m_id, (float*)tmp, n);
        RX_HIP(hipMemcpyAsync(host, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        break;
    case RX_ARRAY_TEMPERATURE:
    case RX_ARRAY_HEAT:
    case RX_ARRAY_PHASE: {
        if (!m_therm) {
            rx_set_error(1, "rxPBDSPH_GPU::GetArray: no thermal state (rxPBDSPH_GPU_SetThermal)");
            return 1;
        }
        k_scatter_out<<<rx_blocks(n, 256), 256, 0, s>>>(m_therm, m_id, tmp, n);
        std::vector<float4> h(n);
        RX_HIP(hipMemcpyAsync(h.data(), tmp, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
        if (!RX_HIP(hipStreamSynchronize(s))) return 1;
        for (uint t = 0; t < n; ++t) host[t] = type == RX_ARRAY_TEMPERATURE ? h[t].x : (type == RX_ARRAY_HEAT ? h[t].y : h[t].z);
        return 0;
    }
    case RX_ARRAY_FEATURES:
        if (count < (int)(16 * n)) {
            rx_set_error(1, "rxPBDSPH_GPU::GetArray: features need 16 floats per particle");
            return 1;
        }
        k_scatter_feat<<<rx_blocks(n, 256), 256, 0, s>>>((const float4*)m_feat, m_id, m_ftmp16, n);
        RX_HIP(hipMemcpyAsync(host, m_ftmp16, sizeof(float4) * 4 * n, hipMemcpyDeviceToHost, s));
        break;
    default:
        rx_set_error(1, "rxPBDSPH_GPU::GetArray: unknown array %d", type);
        return 1;
    }
    return RX_HIP(hipStreamSynchronize(s)) ? 0 : 1;
}

// Thermal state in original particle order: temperature (reference scale, melting at 250), heat, phase (-2 ice,
// -1 melting, 1 freezing, 2 water; NULL: -2 below 250, 2 otherwise).  NULL temps removes the thermal state.
int rxPBDSPH_GPU::SetThermal(const float* temps, const float* heats, const int* phase) {
    drop_graphs();
    if (!m_n) {
        rx_set_error(1, "rxPBDSPH_GPU::SetThermal: set the particles first");
        return 1;
    }
    hipStream_t s = m_stream;
    RX_HIP(hipStreamSynchronize(s));
    if (!temps) {
        for (void** p : {(void**)&m_therm, (void**)&m_therm_t})
            if (*p) {
                RX_HIP(hipFree(*p));
                *p = nullptr;
            }
        return 0;
    }
    if (!m_therm && !(RX_HIP(hipMalloc(&m_therm, sizeof(float4) * m_n)) && RX_HIP(hipMalloc(&m_therm_t, sizeof(float4) * m_n)))) return 1;
    std::vector<float4> h(m_n);
    for (uint i = 0; i < m_n; ++i) {
        int ph = phase ? phase[i] : (temps[i] < 250.0f ? -2 : 2);
        h[i] = make_float4(temps[i], heats ? heats[i] : 0.0f, (float)ph, 0.0f);
    }
    RX_HIP(hipMemcpyAsync(m_t, h.data(), sizeof(float4) * m_n, hipMemcpyHostToDevice, s));
    k_gather_by_id<<<rx_blocks(m_n, 256), 256, 0, s>>>(m_therm, m_t, m_id, m_n);  // into the current slot order
    return RX_HIP(hipStreamSynchronize(s)) ? 0 : 1;
}

rxPBDSPH_GPU::DeviceView rxPBDSPH_GPU::device_view() const {
    return DeviceView{m_x, m_v, m_id, m_xb, m_therm, m_stats, ST_DISP2};
}

int rxPBDSPH_GPU::GetStats(float* out, int count) {
    if (!out || count <= 0) return 1;
    float h[ST_N];
    RX_HIP(hipMemcpyAsync(h, m_stats, sizeof(h), hipMemcpyDeviceToHost, m_stream));
    if (!RX_HIP(hipStreamSynchronize(m_stream))) return 1;
    // uint counters are returned as floats
    h[ST_OVERFLOW] = (float)__builtin_bit_cast(uint, h[ST_OVERFLOW]);
    h[ST_FALLBACK] = (float)__builtin_bit_cast(uint, h[ST_FALLBACK]);
    h[ST_ITERS] = (float)__builtin_bit_cast(uint, h[ST_ITERS]);
    h[ST_BAD] = (float)__builtin_bit_cast(uint, h[ST_BAD]);
    h[ST_MEAN_ABS_START] = h[ST_MEAN_ABS_START] / (float)m_n;
    h[ST_MPOS_START] = h[ST_MPOS_START] / (float)m_n;
    for (int t = 0; t < 16; ++t) h[ST_MPOS_ITER + t] = h[ST_MPOS_ITER + t] / (float)m_n;
    for (int t : {(int)ST_RUN_ITERS, (int)ST_RUN_STEPS, (int)ST_RUN_FAIL, (int)ST_RUN_BAD, (int)ST_MOVERS, (int)ST_MOVERS_PUB,
                  ST_CROSS + 0, ST_CROSS + 1, ST_CROSS + 2, ST_CROSS + 3, (int)ST_LIST_CHUNKS})
        h[t] = (float)__builtin_bit_cast(uint, h[t]);
    h[ST_DISP2] = sqrtf(h[ST_DISP2]);
    h[60] = m_skin / m_env.h;  // host state: the list sizes in use
    h[61] = (float)m_maxn;
    for (int i = 0; i < count && i < ST_N; ++i) out[i] = h[i];
    return 0;
}

int rxPBDSPH_GPU::Sync() { return RX_HIP(hipStreamSynchronize(m_stream)) ? 0 : 1; }

// The solver keeps its own copy of the network's inference weights (made on the solver's stream once the
// network's pending work is done), so the network may be changed or destroyed afterwards.
int rxPBDSPH_GPU::SetNetwork(void* mlp) {
    drop_graphs();
    m_mlp_ready = false;
    if (!mlp) {
        rx_set_error(1, "rxPBDSPH_GPU::SetNetwork: null network");
        return 1;
    }
    if (!m_stream) {
        rx_set_error(1, "rxPBDSPH_GPU::SetNetwork: solver not initialised");
        return 1;
    }
    if (!m_mlp_mem && !RX_HIP(hipMalloc(&m_mlp_mem, rx_mlp_copy_bytes()))) return 1;
    if (!RX_HIP(hipStreamSynchronize(m_stream))) return 1;  // no pending inference reads the old copy
    if (!rx_mlp_copy(mlp, m_mlp_mem, m_stream, &m_mlp)) {
        rx_set_error(1, "rxPBDSPH_GPU::SetNetwork: copying the network failed");
        return 1;
    }
    m_mlp_ready = true;
    return 0;
}

int rxPBDSPH_GPU::begin_mark(int cls) {
    if (!m_profile || !((m_profile_mask >> cls) & 1u)) return -1;
    if (m_event_next + 2 > m_event_pool.size()) {
        for (int t = 0; t < 64; ++t) {
            hipEvent_t e;
            RX_HIP(hipEventCreate(&e));
            m_event_pool.push_back(e);
        }
    }
    Mark m;
    m.cls = cls;
    m.a = m_event_pool[m_event_next++];
    m.b = m_event_pool[m_event_next++];
    RX_HIP(hipEventRecord(m.a, m_stream));
    m_marks.push_back(m);
    return (int)m_marks.size() - 1;
}
void rxPBDSPH_GPU::end_mark(int id) {
    if (id < 0) return;
    RX_HIP(hipEventRecord(m_marks[id].b, m_stream));
    if (m_event_next > 4096) collect_marks();
}
void rxPBDSPH_GPU::collect_marks() {
    if (m_marks.empty()) return;
    RX_HIP(hipStreamSynchronize(m_stream));
    for (const Mark& m : m_marks) {
        float ms = 0.0f;
        RX_HIP(hipEventElapsedTime(&ms, m.a, m.b));
        m_prof_ms[m.cls] += ms;
        m_prof_cnt[m.cls] += 1;
    }
    m_marks.clear();
    m_event_next = 0;
}
int rxPBDSPH_GPU::GetProfile(float* ms, int* count, int n) {
    collect_marks();
    for (int i = 0; i < n && i < 16; ++i) {
        if (ms) ms[i] = (float)m_prof_ms[i];
        if (count) count[i] = (int)m_prof_cnt[i];
        m_prof_ms[i] = 0.0;
        m_prof_cnt[i] = 0;
    }
    return 0;
}

// ============================================================================================ C ABI
extern "C" {

void* rxPBDSPH_GPU_Create(const rxSphEnv* env) {
    if (!env) {
        rx_set_error(1, "rxPBDSPH_GPU_Create: null env");
        return nullptr;
    }
    rxPBDSPH_GPU* p = new rxPBDSPH_GPU();
    if (p->Initialize(*env) != 0) {
        delete p;
        return nullptr;
    }
    return p;
}
void rxPBDSPH_GPU_Destroy(void* h) { delete (rxPBDSPH_GPU*)h; }
int rxPBDSPH_GPU_SetParticles(void* h, const float* pos4, const float* vel4, int n) {
    return h ? ((rxPBDSPH_GPU*)h)->SetParticles(pos4, vel4, n) : 1;
}
int rxPBDSPH_GPU_Update(void* h, float dt, int steps) {
    if (!h) return 1;
    for (int i = 0; i < steps; ++i) {
        const int rc = ((rxPBDSPH_GPU*)h)->Update(dt, i);
        if (rc != 0) return rc;
    }
    return 0;
}
int rxPBDSPH_GPU_GetArray(void* h, int type, float* host, int count) { return h ? ((rxPBDSPH_GPU*)h)->GetArray(type, host, count) : 1; }
int rxPBDSPH_GPU_GetStats(void* h, float* out, int count) { return h ? ((rxPBDSPH_GPU*)h)->GetStats(out, count) : 1; }
int rxPBDSPH_GPU_Sync(void* h) { return h ? ((rxPBDSPH_GPU*)h)->Sync() : 1; }
int rxPBDSPH_GPU_SetOption(void* h, int opt, float value) { return h ? ((rxPBDSPH_GPU*)h)->SetOption(opt, value) : 1; }
int rxPBDSPH_GPU_GetProfile(void* h, float* ms, int* count, int n) { return h ? ((rxPBDSPH_GPU*)h)->GetProfile(ms, count, n) : 1; }
int rxPBDSPH_GPU_SetNetwork(void* h, void* mlp) { return h ? ((rxPBDSPH_GPU*)h)->SetNetwork(mlp) : 1; }
int rxPBDSPH_GPU_SetThermal(void* h, const float* temps, const float* heats, const int* phase) {
    if (!h) return 1;
    return ((rxPBDSPH_GPU*)h)->SetThermal(temps, heats, phase);
}
int rxPBDSPH_GPU_FeatureScales(void* h, float* sd_ref, float* nref) {
    if (!h) return 1;
    ((rxPBDSPH_GPU*)h)->FeatureScales(sd_ref, nref);
    return 0;
}

}  // extern "C"

int rxPBDSPH_GPU::SetOption(int opt, float value) {
    switch (opt) {
    case RX_PBF_OPT_SYNC_EACH_STEP: m_sync_each_step = value != 0.0f; return 0;
    case RX_PBF_OPT_GRAPH:
        m_use_graph = value != 0.0f;
        if (!m_use_graph) drop_graphs();
        return 0;
    case RX_PBF_OPT_RECORD_LAMBDA: return 0;  // the applied multiplier is always accumulated
    case RX_PBF_OPT_RECORD_FEATURES: m_record_features = value != 0.0f; return 0;
    case RX_PBF_OPT_MODE:
        if (value != RX_PBF_MODE_PLAIN && value != RX_PBF_MODE_SCALED && value != RX_PBF_MODE_LEARNED) {
            rx_set_error(1, "rxPBDSPH_GPU::SetOption: unknown mode %g", value);
            return 1;
        }
        m_mode = (int)value;
        return 0;
    case RX_PBF_OPT_ADAPTIVE: m_adaptive = value != 0.0f; return 0;
    case RX_PBF_OPT_WARM_SCALE: m_warm_scale = value; return 0;
    case RX_PBF_OPT_TOL: m_tol = value; return 0;
    case RX_PBF_OPT_JUDGE:
        if (value != 0.0f && value != 1.0f) {
            rx_set_error(1, "rxPBDSPH_GPU::SetOption: judge must be 0 (max) or 1 (mean)");
            return 1;
        }
        m_judge = (int)value;
        return 0;
    case RX_PBF_OPT_RESET_RUN:
        if (m_stats) k_fill<float><<<1, 64, 0, m_stream>>>(m_stats + ST_RUN, 0.0f, ST_N - ST_RUN);
        return RX_HIP(hipGetLastError()) ? 0 : 1;
    case RX_PBF_OPT_ITERATIONS: m_env.max_iter = m_env.min_iter = (int)value; return 0;
    case RX_PBF_OPT_PROFILE:
        m_profile = value != 0.0f;
        return 0;
    case RX_PBF_OPT_PROFILE_MASK:
        m_profile_mask = (unsigned)value;
        return 0;
    case RX_PBF_OPT_HEAT_TD: m_heat_td = value; return 0;
    case RX_PBF_OPT_MAX_NEIGHBORS: {
        int m = (int)value;
        if (m < 8 || m > 512 || (m & 3)) {
            rx_set_error(1, "rxPBDSPH_GPU::SetOption: list capacity must be a multiple of 4 in 8..512");
            return 1;
        }
        if (m != m_maxn && m_list) {
            uint* nl = nullptr;  // the new list first: on failure the old one stays in use
            if (!RX_HIP(hipMalloc(&nl, list_bytes(m_n, m)))) return 1;
            drop_graphs();
            RX_HIP(hipStreamSynchronize(m_stream));
            RX_HIP(hipFree(m_list));
            m_list = nl;
        }
        m_maxn = m_maxn_target = m;
        m_maxn_cap = std::max(512, m);
        m_have_list = false;
        return 0;
    }
    case RX_PBF_OPT_ADAPT_SKIN:
        m_adapt = value != 0.0f;
        if (!m_adapt) m_level_target = 0;  // back to the set skin at the next step
        return 0;
    case RX_PBF_OPT_HEAT_DT: m_heat_dt = value; return 0;
    case RX_PBF_OPT_WARM_CLAMP:
    case RX_PBF_OPT_DEFER:
        m_defer_on = value != 0.0f ? 1 : 0;
        drop_graphs();
        return 0;
    case RX_PBF_OPT_FIRST_DMAX:
        if (!(value >= 0.0f)) {
            rx_set_error(1, "rxPBDSPH_GPU::SetOption: the bound must be >= 0 (0 = none)");
            return 1;
        }
        (opt == RX_PBF_OPT_WARM_CLAMP ? m_warm_kappa : m_first_dmax) = value;
        drop_graphs();
        return 0;
    case RX_PBF_OPT_GRAD_MASS:
        m_grad_mass = value != 0.0f;
        lattice_refs();  // the feature scales follow the gradient
        drop_graphs();
        return 0;
    case RX_PBF_OPT_LIST_TILE:
        if (value < 0.0f || value > 4.0f) {
            rx_set_error(1, "rxPBDSPH_GPU::SetOption: list tile variant must be 0..4");
            return 1;
        }
        m_list_tile = (int)value;
        return 0;
    case RX_PBF_OPT_SKIN:
        if (value < 0.0f) return 1;
        m_skin = m_skin_base = value * m_env.h;
        m_level = m_level_target = 0;
        m_calm = 0;
        m_grid.setup(m_params, m_env.h + m_skin);
        m_have_list = false;
        if (!m_grid.reserve(m_n)) {
            m_n = 0;  // Update refuses until a SetParticles succeeds
            return 1;
        }
        return 0;
    default:
        rx_set_error(1, "rxPBDSPH_GPU::SetOption: unknown option %d", opt);
        return 1;
    }
}

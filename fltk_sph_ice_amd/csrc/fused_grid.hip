// fused_grid.hip -- stable counting-sort rebuild of the fused solvers' neighbour grid.
#include "fused_grid.h"

namespace {

// Hash + count.  Lanes of a wavefront hold consecutive slots, which after the previous rebuild mostly share
// a cell: each run of equal hashes takes one atomic for the whole run (run heads found with a ballot).
__global__ __launch_bounds__(256) void k_fg_hash_count(const float4* __restrict__ pos, uint n, FGridK g, uint* __restrict__ hash,
                                                       uint* __restrict__ slot, uint* __restrict__ cnt) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < n;
    uint h = valid ? fg_hash(fg_pos(f3(pos[i]), g), g) : 0xFFFFFFFFu;
    const int lane = threadIdx.x & 63;
    uint prev = __shfl_up(h, 1, 64);
    bool head = valid && (lane == 0 || prev != h);
    unsigned long long heads = __ballot(head);
    unsigned long long validm = __ballot(valid);
    if (!valid) return;
    // my run's head: highest head lane <= lane
    unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int my_head = 63 - __clzll(heads & le);
    // next head after me (or the end of the valid lanes)
    unsigned long long gt = heads & ~le;
    int next = gt ? (__ffsll((long long)gt) - 1) : (64 - __clzll(validm));
    uint base = 0;
    if (head) base = atomicAdd(&cnt[h], (uint)(next - lane));
    base = __shfl(base, my_head, 64);
    hash[i] = h;
    slot[i] = base + (uint)(lane - my_head);
}

__global__ void k_fg_scatter(const uint* __restrict__ hash, const uint* __restrict__ slot, const uint* __restrict__ off, uint n,
                             uint* __restrict__ tsrc, uint* __restrict__ thash) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint h = hash[i];
    uint p = off[h] + slot[i];
    tsrc[p] = i;
    thash[p] = h;
}

// Final place of each element of a cell = count of the cell's elements with a smaller previous slot.
__global__ void k_fg_rank(const uint* __restrict__ tsrc, const uint* __restrict__ thash, const uint* __restrict__ off, uint n,
                          uint* __restrict__ perm, uint* __restrict__ hash_out) {
    uint p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint h = thash[p];
    uint s = off[h], e = off[h + 1];
    uint src = tsrc[p];
    uint r = 0;
    for (uint q = s; q < e; ++q) r += (tsrc[q] < src) ? 1u : 0u;
    perm[s + r] = src;
    hash_out[s + r] = h;
}

}  // namespace

void FusedGrid::setup(const rxSimParams& p, float cell) {
    k.origin = p.WorldOrigin;
    k.cell = cell;
    k.inv_cell = 1.0f / cell;
    float3 ws = f3(p.WorldMax.x - p.WorldOrigin.x, p.WorldMax.y - p.WorldOrigin.y, p.WorldMax.z - p.WorldOrigin.z);
    k.gx = (int)(ws.x / cell) + 1;
    k.gy = (int)(ws.y / cell) + 1;
    k.gz = (int)(ws.z / cell) + 1;
    k.ncells = (uint)k.gx * (uint)k.gy * (uint)k.gz;
}

bool FusedGrid::reserve(uint np) {
    bool ok = true;
    if (np > cap_n) {
        for (uint** p : {&hash, &perm, &tmp_hash, &slot, &tsrc})
            if (*p) {
                RX_HIP(hipFree(*p));
                *p = nullptr;
            }
        cap_n = 0;
        ok = RX_HIP(hipMalloc(&hash, sizeof(uint) * np)) && RX_HIP(hipMalloc(&perm, sizeof(uint) * np)) &&
             RX_HIP(hipMalloc(&tmp_hash, sizeof(uint) * np)) && RX_HIP(hipMalloc(&slot, sizeof(uint) * np)) &&
             RX_HIP(hipMalloc(&tsrc, sizeof(uint) * np));
        if (ok) cap_n = np;
    }
    if (ok && (size_t)k.ncells + 1 > cap_c) {
        if (off) RX_HIP(hipFree(off));
        if (scan_ws) RX_HIP(hipFree(scan_ws));
        off = nullptr;
        scan_ws = nullptr;
        cap_c = 0;
        ok = RX_HIP(hipMalloc(&off, sizeof(uint) * ((size_t)k.ncells + 2))) &&
             RX_HIP(hipMalloc(&scan_ws, rx_scan_ws_bytes((size_t)k.ncells + 1)));
        if (ok) cap_c = (size_t)k.ncells + 1;
    }
    if (!ok) release();
    return ok;
}

void FusedGrid::release() {
    for (uint** p : {&hash, &off, &perm, &tmp_hash, &slot, &tsrc})
        if (*p) {
            RX_HIP(hipFree(*p));
            *p = nullptr;
        }
    if (scan_ws) {
        RX_HIP(hipFree(scan_ws));
        scan_ws = nullptr;
    }
    cap_n = cap_c = 0;
}

void FusedGrid::build(const float4* pos, hipStream_t s) {
    if (!n) return;
    rx_fill<uint>(off, 0u, (size_t)k.ncells + 1, s);
    unsigned b = rx_blocks(n, 256);
    k_fg_hash_count<<<b, 256, 0, s>>>(pos, n, k, tmp_hash, slot, off);
    rx_scan_u32(off, off, k.ncells, true, s, scan_ws);  // off[c] = first slot of cell c, off[ncells] = n
    k_fg_scatter<<<b, 256, 0, s>>>(tmp_hash, slot, off, n, tsrc, perm /* as scratch for the cell ids */);
    // perm doubles as the cell-id scratch of the scatter; the rank pass reads it and writes hash/perm anew
    k_fg_rank<<<b, 256, 0, s>>>(tsrc, perm, off, n, slot /* perm out */, hash);
    std::swap(perm, slot);
}

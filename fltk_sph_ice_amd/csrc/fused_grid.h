// fused_grid.h -- the neighbour grid of the fused solvers.
//
// Particle state is kept permanently in cell order.  Each rebuild hashes the current positions and radix-sorts
// (cell, previous slot) pairs, which makes the permutation stable (ties in previous-slot order) and
// deterministic in linear time; the cell offsets (off[c] = first slot of cell c, off[G] = N, so empty cells need
// no sentinel and a row of cells is one slot range) are found by binary search in the sorted cells.
// `perm[new] = old` drives the gather of every state array (optionally done by the sort's last pass).
#pragma once

#include "rx_common.h"

struct FGridK {
    float3 origin;
    float inv_cell;  // unused by the hash (which divides), kept for callers that bound cell ranges
    float cell;
    int gx, gy, gz;
    uint ncells;
};

__device__ inline int3 fg_pos(float3 p, const FGridK& g) {
    int3 c;
    c.x = (int)floorf((p.x - g.origin.x) / g.cell);
    c.y = (int)floorf((p.y - g.origin.y) / g.cell);
    c.z = (int)floorf((p.z - g.origin.z) / g.cell);
    c.x = min(max(c.x, 0), g.gx - 1);
    c.y = min(max(c.y, 0), g.gy - 1);
    c.z = min(max(c.z, 0), g.gz - 1);
    return c;
}
__device__ inline uint fg_hash(int3 c, const FGridK& g) { return (uint)((c.z * g.gy + c.y) * g.gx + c.x); }
__device__ inline int3 fg_unhash(uint h, const FGridK& g) {
    int3 c;
    c.x = (int)(h % (uint)g.gx);
    uint r = h / (uint)g.gx;
    c.y = (int)(r % (uint)g.gy);
    c.z = (int)(r / (uint)g.gy);
    return c;
}

// Visits every slot j of the cells within `rad` cells of cell c (clamped to the grid): (2 rad + 1)^2 rows,
// each one contiguous slot range.
template <class F>
__device__ inline void fg_for_stencil(int3 c, const FGridK& g, const uint* __restrict__ off, F&& f, int rad = 1) {
    int x0 = max(c.x - rad, 0), x1 = min(c.x + rad, g.gx - 1);
    int y0 = max(c.y - rad, 0), y1 = min(c.y + rad, g.gy - 1);
    int z0 = max(c.z - rad, 0), z1 = min(c.z + rad, g.gz - 1);
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y) {
            uint hb = (uint)((z * g.gy + y) * g.gx);
            uint s = off[hb + x0], e = off[hb + x1 + 1];
            for (uint j = s; j < e; ++j) f(j);
        }
}

struct FusedGrid {
    FGridK k{};
    uint n = 0;
    // device arrays
    uint* hash = nullptr;    // cell of each slot, in the new (sorted) order after build
    uint* off = nullptr;     // ncells + 1 offsets
    uint* perm = nullptr;    // new slot -> old slot
    uint* tmp_hash = nullptr;  // cells in the current order
    uint* slot = nullptr;      // sort ping-pong pair
    uint* tsrc = nullptr;
    void* radix_ws = nullptr;  // own sort workspace, so that a rebuild can be captured into a graph
    size_t cap_n = 0, cap_c = 0;

    void setup(const rxSimParams& p, float cell);
    bool reserve(uint n_particles);  // false (and every array released) when an allocation fails
    void release();
    // Builds the grid on pos (in the current slot order); afterwards hash/off describe the new order and
    // perm maps new slots to current slots.  gather (may be null) names state arrays to reorder on the way.
    void build(const float4* pos, hipStream_t s, const struct RadixGather* gather = nullptr);
};

template <class T> __global__ void k_gather(T* __restrict__ dst, const T* __restrict__ src, const uint* __restrict__ perm, uint n) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}
template <class T> inline void fg_gather(T* dst, const T* src, const uint* perm, uint n, hipStream_t s) {
    if (n) k_gather<T><<<rx_blocks(n, 256), 256, 0, s>>>(dst, src, perm, n);
}

// fused_grid.hip -- buffers of the fused solvers' neighbour grid; the rebuild itself is the radix sort of
// radix.hip.
#include "fused_grid.h"
#include "radix.h"

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
        if (radix_ws) RX_HIP(hipFree(radix_ws));
        radix_ws = nullptr;
        cap_n = 0;
        ok = RX_HIP(hipMalloc(&hash, sizeof(uint) * np)) && RX_HIP(hipMalloc(&perm, sizeof(uint) * np)) &&
             RX_HIP(hipMalloc(&tmp_hash, sizeof(uint) * np)) && RX_HIP(hipMalloc(&slot, sizeof(uint) * np)) &&
             RX_HIP(hipMalloc(&tsrc, sizeof(uint) * np)) && RX_HIP(hipMalloc(&radix_ws, rx_radix_ws_bytes(np)));
        if (ok) cap_n = np;
    }
    if (ok && (size_t)k.ncells + 1 > cap_c) {
        if (off) RX_HIP(hipFree(off));
        off = nullptr;
        cap_c = 0;
        ok = RX_HIP(hipMalloc(&off, sizeof(uint) * ((size_t)k.ncells + 2)));
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
    if (radix_ws) {
        RX_HIP(hipFree(radix_ws));
        radix_ws = nullptr;
    }
    cap_n = cap_c = 0;
}

void FusedGrid::build(const float4* pos, hipStream_t s, const RadixGather* gather) {
    if (!n) return;
    rx_grid_sort_gather(pos, k, n, tmp_hash, hash, perm, slot, tsrc, off, radix_ws, gather ? *gather : RadixGather{}, s);
}

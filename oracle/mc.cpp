// mc.cpp -- CPU restatement of the reference's marching-cubes stages (rx_mc.cu:148-389, rx_mc_kernel.cu:319-636).
// TEST INFRASTRUCTURE ONLY: the oracle of fltk_sph_ice_amd/csrc/mc.hip, loaded by tests/ through liboracle.so.
//
// Pinned by: the case tables' digests against the reference's tables (tests/golden/mc_tables_digest.json), the
// closed-surface invariants of a sphere and a torus (every mesh edge shared by two triangles with opposite
// orientations, Euler characteristic 2 / 0), vertices of a linear field on its exact iso-plane, and hand-worked
// single-voxel cases (tests/test_oracle_mc.py).
//
// Sequential loops in the order the kernels define; every multiply and add rounds on its own (the library
// compiles these steps without contraction as well).  The case tables are the library's own
// (mc_tables.h), pinned by the digests above.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../fltk_sph_ice_amd/csrc/mc_tables.h"

namespace {

struct U3 {
    uint32_t x, y, z;
};

U3 unflatten(uint32_t i, U3 n) {
    const uint32_t w = i % (n.x * n.y);
    return {w % n.x, w / n.x, i / (n.x * n.y)};
}
// rx_mc_kernel.cu:153-160 (calcGridIdx3) and 87-95 (sampleVolume2): coordinates clamped to the last index
uint32_t flat_clamped(U3 p, U3 n) {
    p.x = p.x < n.x - 1 ? p.x : n.x - 1;
    p.y = p.y < n.y - 1 ? p.y : n.y - 1;
    p.z = p.z < n.z - 1 ? p.z : n.z - 1;
    return (p.z * n.y + p.y) * n.x + p.x;
}
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = s;
        s += v;
    }
}

}  // namespace

extern "C" {

// The case tables as compiled (edge masks, vertex counts, triangle edges with 0xFF after the last).
void orc_mc_tables(uint16_t* edge, uint8_t* nverts, uint8_t* tri) {
    for (int c = 0; c < 256; ++c) {
        edge[c] = mc::kTables.edge_mask[c];
        nverts[c] = mc::kTables.nverts[c];
        for (int k = 0; k < 16; ++k) tri[16 * c + k] = mc::kTables.tri[c][k];
    }
}

// rx_mc.cu:148-231 with ClassifyVoxel2 (rx_mc_kernel.cu:319-392) and CompactVoxels (:463-472).
// out3 = (active voxels, sum of vertex counts, sum of triangle counts).
void orc_mc_tri_num(const float* vol, const uint32_t* n3, uint32_t nvox, float thr, uint32_t* voxbit, uint32_t* vnum,
                    uint32_t* vscan, uint32_t* tnum, uint32_t* tscan, uint32_t* occ, uint32_t* occscan, uint32_t* compacted,
                    uint32_t* out3) {
    const U3 n{n3[0], n3[1], n3[2]};
    static const uint32_t off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (uint32_t i = 0; i < nvox; ++i) {
        const U3 g = unflatten(i, n);
        uint32_t c = 0;
        for (int k = 0; k < 8; ++k) {
            const float f = vol[flat_clamped({g.x + off[k][0], g.y + off[k][1], g.z + off[k][2]}, n)];
            if (f < thr) c += 1u << k;
        }
        const uint32_t nv = mc::kTables.nverts[c];
        voxbit[i] = c;
        vnum[i] = nv;
        tnum[i] = nv / 3;
        occ[i] = nv > 0;
    }
    out3[0] = out3[1] = out3[2] = 0;
    exclusive_scan(occ, occscan, nvox);
    const uint32_t active = nvox ? occscan[nvox - 1] + occ[nvox - 1] : 0;
    out3[0] = active;
    if (!active) return;
    for (uint32_t i = 0; i < nvox; ++i)
        if (occ[i]) compacted[occscan[i]] = i;
    exclusive_scan(vnum, vscan, nvox);
    exclusive_scan(tnum, tscan, nvox);
    out3[1] = vscan[nvox - 1] + vnum[nvox - 1];
    out3[2] = tscan[nvox - 1] + tnum[nvox - 1];
}

// rx_mc.cu:252-315 with CalVertexEdge (rx_mc_kernel.cu:395-460), vertexInterp (:231-237) and CompactEdges
// (:483-507).  esize9 = edge_size[3] as 9 uints, nedge4 = num_edge[4].  Returns the vertex count.
uint32_t orc_mc_edge_vrts(const float* vol, const uint32_t* esize9, const uint32_t* nedge4, const uint32_t* n3, const float* h3,
                          const float* lo3, float thr, float* edge_vrts4, float* compacted4, uint32_t* eocc, uint32_t* escan) {
    const U3 n{n3[0], n3[1], n3[2]};
    uint32_t pos = 0;
    for (int d = 0; d < 3; ++d) {
        const U3 es{esize9[3 * d], esize9[3 * d + 1], esize9[3 * d + 2]};
        const uint32_t dx = d == 0, dy = d == 1, dz = d == 2;
        for (uint32_t j = 0; j < nedge4[d]; ++j) {
            const uint32_t i = pos + j;
            const U3 g = unflatten(j, es);
            const float f0 = vol[flat_clamped(g, n)];
            const float f1 = vol[flat_clamped({g.x + dx, g.y + dy, g.z + dz}, n)];
            if ((f0 < thr) != (f1 < thr)) {
                float p0[3], p1[3];
                p0[0] = lo3[0] + (float)g.x * h3[0];
                p0[1] = lo3[1] + (float)g.y * h3[1];
                p0[2] = lo3[2] + (float)g.z * h3[2];
                p1[0] = p0[0] + (float)dx * h3[0];
                p1[1] = p0[1] + (float)dy * h3[1];
                p1[2] = p0[2] + (float)dz * h3[2];
                const float t = (thr - f0) / (f1 - f0);
                for (int a = 0; a < 3; ++a) edge_vrts4[4 * (size_t)i + a] = p0[a] + t * (p1[a] - p0[a]);
                edge_vrts4[4 * (size_t)i + 3] = 1.0f;
                eocc[i] = 1;
            } else {
                eocc[i] = 0;
            }
        }
        pos += nedge4[d];
    }
    const uint32_t total = nedge4[3];
    exclusive_scan(eocc, escan, total);
    const uint32_t nv = total ? escan[total - 1] + eocc[total - 1] : 0;
    for (uint32_t i = 0; i < total; ++i)
        if (eocc[i]) std::memcpy(compacted4 + 4 * (size_t)escan[i], edge_vrts4 + 4 * (size_t)i, 16);
    return nv;
}

// rx_mc.cu:332-356 with GenerateTriangles3 (rx_mc_kernel.cu:510-591).
void orc_mc_tri(uint32_t* tris3, const uint32_t* voxbit, const uint32_t* tscan, const uint32_t* compacted, const uint32_t* escan,
                const uint32_t* esize9, const uint32_t* nedge4, const uint32_t* n3, uint32_t active, uint32_t nvrts, uint32_t ntris) {
    const U3 n{n3[0], n3[1], n3[2]};
    const U3 ex{esize9[0], esize9[1], esize9[2]}, ey{esize9[3], esize9[4], esize9[5]}, ez{esize9[6], esize9[7], esize9[8]};
    const uint32_t oy = nedge4[0], oz = nedge4[0] + nedge4[1];
    for (uint32_t idx = 0; idx < active; ++idx) {
        const uint32_t voxel = compacted[idx];
        const U3 g = unflatten(voxel, n);
        const uint32_t c = voxbit[voxel];
        uint32_t v[12];
        v[0] = escan[flat_clamped({g.x, g.y, g.z}, ex)];
        v[1] = escan[flat_clamped({g.x + 1, g.y, g.z}, ey) + oy];
        v[2] = escan[flat_clamped({g.x, g.y + 1, g.z}, ex)];
        v[3] = escan[flat_clamped({g.x, g.y, g.z}, ey) + oy];
        v[4] = escan[flat_clamped({g.x, g.y, g.z + 1}, ex)];
        v[5] = escan[flat_clamped({g.x + 1, g.y, g.z + 1}, ey) + oy];
        v[6] = escan[flat_clamped({g.x, g.y + 1, g.z + 1}, ex)];
        v[7] = escan[flat_clamped({g.x, g.y, g.z + 1}, ey) + oy];
        v[8] = escan[flat_clamped({g.x, g.y, g.z}, ez) + oz];
        v[9] = escan[flat_clamped({g.x + 1, g.y, g.z}, ez) + oz];
        v[10] = escan[flat_clamped({g.x + 1, g.y + 1, g.z}, ez) + oz];
        v[11] = escan[flat_clamped({g.x, g.y + 1, g.z}, ez) + oz];
        const uint32_t nt = mc::kTables.nverts[c] / 3;
        for (uint32_t t = 0; t < nt; ++t) {
            const uint32_t index = tscan[voxel] + t;
            uint32_t a = v[mc::kTables.tri[c][3 * t]], b = v[mc::kTables.tri[c][3 * t + 1]], d = v[mc::kTables.tri[c][3 * t + 2]];
            a = a < nvrts - 1 ? a : nvrts - 1;
            b = b < nvrts - 1 ? b : nvrts - 1;
            d = d < nvrts - 1 ? d : nvrts - 1;
            if (index < ntris) {
                tris3[3 * (size_t)index + 0] = b;
                tris3[3 * (size_t)index + 1] = d;
                tris3[3 * (size_t)index + 2] = a;
            }
        }
    }
}

// rx_mc.cu:366-389 with CalVertexNormalKernel (rx_mc_kernel.cu:594-626) and NormalizeKernel (:628-636): the
// triangle normals accumulated in triangle order, then n / |n| (zero sums stay zero).
void orc_mc_nrm(float* nrm3, const uint32_t* tris3, const float* vrts4, uint32_t nvrts, uint32_t ntris) {
    std::memset(nrm3, 0, sizeof(float) * 3 * (size_t)nvrts);
    for (uint32_t t = 0; t < ntris; ++t) {
        const uint32_t id[3] = {tris3[3 * (size_t)t], tris3[3 * (size_t)t + 1], tris3[3 * (size_t)t + 2]};
        const float* a = vrts4 + 4 * (size_t)id[0];
        const float* b = vrts4 + 4 * (size_t)id[1];
        const float* c = vrts4 + 4 * (size_t)id[2];
        const float e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const float nn[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
        for (int k = 0; k < 3; ++k) {
            if (id[k] >= nvrts) continue;
            float* o = nrm3 + 3 * (size_t)id[k];
            o[0] = o[0] + nn[0];
            o[1] = o[1] + nn[1];
            o[2] = o[2] + nn[2];
        }
    }
    for (uint32_t v = 0; v < nvrts; ++v) {
        float* o = nrm3 + 3 * (size_t)v;
        const float l = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        if (l > 0.0f) {
            o[0] = o[0] / l;
            o[1] = o[1] / l;
            o[2] = o[2] / l;
        }
    }
}

}  // extern "C"

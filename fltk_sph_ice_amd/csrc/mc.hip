// mc.hip -- marching-cubes mesh extraction from a density lattice: the RX_CUMC_USE_GEOMETRY path of the ABI
// (rx_cu_funcs.cuh:184-212, rx_mc.cu:135-344, rx_mc_kernel.cu:319-640).
//
// Sequence (as rxMCMeshGPU::CreateMeshV, rx_mc_gpu.cpp:336-366, calls it):
//   CuMCCalTriNum    classify every voxel (case index, vertex and triangle counts), scan and compact the voxels
//                    that hold a surface piece, scan the counts; returns active voxels, 3 x triangles, triangles
//   CuMCCalEdgeVrts  one vertex per lattice edge that the iso level crosses, scanned and compacted
//   CuMCCalTri       the triangles of every active voxel as indices into the compacted vertices
//   CuMCCalNrm       vertex normals: sum of the incident triangles' cross products, normalised
//
// Lattice.  The volume holds nx*ny*nz samples, x fastest.  Voxel (x, y, z) has its corner k at (x, y, z) + the
// offset of corner k (mc_tables.h); a corner index past the last sample reads the last sample (rx_mc_kernel.cu:
// 87-95), so the outermost voxel layers are degenerate in the clamped direction.  Edges of direction d are
// numbered over edge_size[d] = grid + 1 in the two other directions; the three direction blocks are
// concatenated (x, y, z) in one array of num_edge[3] entries.
//
// Rounding.  Vertex positions, triangle normals and their sums are computed with contraction switched off, so
// each multiply and add rounds on its own as in the CPU restatement; with IEEE division and square root
// (hipcc's default) the vertices and normals are bit-identical to it.  The normal sums are taken in triangle
// order per vertex (an incidence list built by counting sort, each vertex's list ordered before summing), so
// the result does not depend on atomic ordering.
#include "rx_common.h"
#include "mc_tables.h"
#include "../../include/rx_cu_mesh.h"

namespace {

constexpr int kTpb = 256;

__constant__ mc::Tables c_mc = mc::kTables;

__device__ inline uint3 unflatten(uint i, uint3 n) {
    const uint w = i % (n.x * n.y);
    return make_uint3(w % n.x, w / n.x, i / (n.x * n.y));
}
__device__ inline uint flatten_clamped(uint3 p, uint3 n) {
    p.x = min(p.x, n.x - 1);
    p.y = min(p.y, n.y - 1);
    p.z = min(p.z, n.z - 1);
    return (p.z * n.y + p.y) * n.x + p.x;
}
__device__ inline float sample(const float* __restrict__ v, uint3 p, uint3 n) { return v[flatten_clamped(p, n)]; }

// rx_mc_kernel.cu:319-392 (ClassifyVoxel2): corner k below the threshold sets bit k of the case index.
__global__ __launch_bounds__(kTpb) void k_mc_classify(const float* __restrict__ vol, uint3 n, uint nvox, float thr,
                                                      uint* __restrict__ voxbit, uint* __restrict__ vnum,
                                                      uint* __restrict__ tnum, uint* __restrict__ occ) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nvox) return;
    const uint3 g = unflatten(i, n);
    const uint x1 = g.x + 1, y1 = g.y + 1, z1 = g.z + 1;
    uint c = 0;
    c |= (uint)(sample(vol, make_uint3(g.x, g.y, g.z), n) < thr);
    c |= (uint)(sample(vol, make_uint3(x1, g.y, g.z), n) < thr) << 1;
    c |= (uint)(sample(vol, make_uint3(x1, y1, g.z), n) < thr) << 2;
    c |= (uint)(sample(vol, make_uint3(g.x, y1, g.z), n) < thr) << 3;
    c |= (uint)(sample(vol, make_uint3(g.x, g.y, z1), n) < thr) << 4;
    c |= (uint)(sample(vol, make_uint3(x1, g.y, z1), n) < thr) << 5;
    c |= (uint)(sample(vol, make_uint3(x1, y1, z1), n) < thr) << 6;
    c |= (uint)(sample(vol, make_uint3(g.x, y1, z1), n) < thr) << 7;
    const uint nv = c_mc.nverts[c];
    voxbit[i] = c;
    vnum[i] = nv;
    tnum[i] = nv / 3u;
    occ[i] = nv > 0u;
}

// rx_mc_kernel.cu:463-507 (CompactVoxels / CompactEdges)
__global__ __launch_bounds__(kTpb) void k_compact_index(uint* __restrict__ out, const uint* __restrict__ occ,
                                                        const uint* __restrict__ scan, uint n) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n && occ[i]) out[scan[i]] = i;
}
__global__ __launch_bounds__(kTpb) void k_compact_f4(float4* __restrict__ out, const uint* __restrict__ occ,
                                                     const uint* __restrict__ scan, const float4* __restrict__ in, uint n) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i < n && occ[i]) out[scan[i]] = in[i];
}

// total of an exclusive scan: scan[n-1] + in[n-1], for up to three arrays at once
__global__ void k_scan_totals(const uint* a_in, const uint* a_scan, const uint* b_in, const uint* b_scan,
                              const uint* c_in, const uint* c_scan, uint n, uint* __restrict__ out) {
    if (threadIdx.x != 0) return;
    out[0] = a_in ? a_scan[n - 1] + a_in[n - 1] : 0u;
    out[1] = b_in ? b_scan[n - 1] + b_in[n - 1] : 0u;
    out[2] = c_in ? c_scan[n - 1] + c_in[n - 1] : 0u;
}

struct EdgeK {
    uint3 size[3];  // edges per axis, per direction
    uint first[3];  // index of the first edge of each direction
    uint total;
    uint3 n;        // samples per axis
    float3 h, lo;
    float thr;
};

#pragma clang fp contract(off)
// rx_mc_kernel.cu:395-460 (CalVertexEdge) with vertexInterp (rx_mc_kernel.cu:231-237): lerp(p0, p1, t) =
// p0 + t (p1 - p0), t = (iso - f0) / (f1 - f0), p1 = p0 + the voxel width along the edge.
__global__ __launch_bounds__(kTpb) void k_mc_edges(const float* __restrict__ vol, EdgeK k, float4* __restrict__ vrts,
                                                   uint* __restrict__ occ) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= k.total) return;
    const uint d = i >= k.first[2] ? 2u : (i >= k.first[1] ? 1u : 0u);
    const uint j = i - k.first[d];
    const uint3 g = unflatten(j, k.size[d]);
    const uint3 dir = make_uint3(d == 0u, d == 1u, d == 2u);
    const float f0 = sample(vol, g, k.n);
    const float f1 = sample(vol, make_uint3(g.x + dir.x, g.y + dir.y, g.z + dir.z), k.n);
    const bool b0 = f0 < k.thr, b1 = f1 < k.thr;
    if (b0 != b1) {
        float3 p0, p1;
        p0.x = k.lo.x + (float)g.x * k.h.x;
        p0.y = k.lo.y + (float)g.y * k.h.y;
        p0.z = k.lo.z + (float)g.z * k.h.z;
        p1.x = p0.x + (float)dir.x * k.h.x;
        p1.y = p0.y + (float)dir.y * k.h.y;
        p1.z = p0.z + (float)dir.z * k.h.z;
        const float t = (k.thr - f0) / (f1 - f0);
        vrts[i] = make_float4(p0.x + t * (p1.x - p0.x), p0.y + t * (p1.y - p0.y), p0.z + t * (p1.z - p0.z), 1.0f);
        occ[i] = 1u;
    } else {
        occ[i] = 0u;
    }
}

// rx_mc_kernel.cu:594-626 (CalVertexNormalKernel): n = (v1 - v0) x (v2 - v0), unnormalised
__device__ inline float3 tri_normal(const float4* __restrict__ v, uint3 t) {
    const float4 a = v[t.x], b = v[t.y], c = v[t.z];
    const float3 e0 = make_float3(b.x - a.x, b.y - a.y, b.z - a.z);
    const float3 e1 = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    return make_float3(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x);
}

// Per vertex: sum of the normals of its incident triangles in ascending triangle order, then normalised
// (rx_mc_kernel.cu:628-636: v / |v|, left as is when |v| = 0).  The list of vertex v is inc[start[v] ..
// start[v + 1]); its order is chosen by the fill kernel's atomics, so the walk picks the entries in ascending
// order (a selection pass per entry: lists hold at most a few tens of triangles).
__global__ __launch_bounds__(kTpb) void k_mc_vertex_normals(const float4* __restrict__ vrts, const uint3* __restrict__ tris,
                                                            const uint* __restrict__ start, const uint* __restrict__ inc,
                                                            float* __restrict__ nrm, uint nv) {
    const uint v = blockIdx.x * kTpb + threadIdx.x;
    if (v >= nv) return;
    const uint s = start[v], e = start[v + 1];
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    uint last = 0u;
    bool first = true;
    for (uint k = s; k < e; ++k) {
        uint best = 0xFFFFFFFFu;
        for (uint q = s; q < e; ++q) {
            const uint t = inc[q];
            if ((first || t > last) && t < best) best = t;
        }
        // a triangle listed twice for this vertex (repeated index) is taken twice
        uint mult = 0u;
        for (uint q = s; q < e; ++q) mult += inc[q] == best;
        const float3 tn = tri_normal(vrts, tris[best]);
        for (uint m = 0; m < mult; ++m) acc = make_float3(acc.x + tn.x, acc.y + tn.y, acc.z + tn.z);
        k += mult - 1u;
        last = best;
        first = false;
    }
    const float l = sqrtf(acc.x * acc.x + acc.y * acc.y + acc.z * acc.z);
    if (l > 0.0f) acc = make_float3(acc.x / l, acc.y / l, acc.z / l);
    nrm[3 * (size_t)v + 0] = acc.x;
    nrm[3 * (size_t)v + 1] = acc.y;
    nrm[3 * (size_t)v + 2] = acc.z;
}
#pragma clang fp contract(on)

struct TriK {
    uint3 esize[3];
    uint3 ecount;  // edges of the x and y direction blocks (offsets of the y and z blocks)
    uint3 n;
    uint active;
    uint maxv;     // vertex indices are clamped to maxv - 1
    uint ntris;
};

// rx_mc_kernel.cu:510-591 (GenerateTriangles3): the edge-vertex index of each of the voxel's 12 edges is the
// exclusive scan of the edge occupancy at that edge; triangle i of the case, edges (e0, e1, e2), is written at
// voxelTrisScan + i as (v[e1], v[e2], v[e0]).
__global__ __launch_bounds__(kTpb) void k_mc_triangles(uint3* __restrict__ tris, const uint* __restrict__ tscan,
                                                       const uint* __restrict__ escan, const uint* __restrict__ voxbit,
                                                       const uint* __restrict__ compacted, TriK k) {
    const uint idx = blockIdx.x * kTpb + threadIdx.x;
    if (idx >= k.active) return;
    const uint voxel = compacted[idx];
    const uint3 g = unflatten(voxel, k.n);
    const uint c = voxbit[voxel];
    const uint ox = 0u, oy = k.ecount.x, oz = k.ecount.x + k.ecount.y;
    const uint x1 = g.x + 1, y1 = g.y + 1, z1 = g.z + 1;
    // edge -> lattice edge (corner numbering of mc_tables.h)
    uint ev[12];
    const uint mask = c_mc.edge_mask[c];
    auto at = [&](int e, uint3 p, int d, uint off) {
        if (mask & (1u << e)) ev[e] = escan[flatten_clamped(p, k.esize[d]) + off];
    };
    at(0, make_uint3(g.x, g.y, g.z), 0, ox);
    at(1, make_uint3(x1, g.y, g.z), 1, oy);
    at(2, make_uint3(g.x, y1, g.z), 0, ox);
    at(3, make_uint3(g.x, g.y, g.z), 1, oy);
    at(4, make_uint3(g.x, g.y, z1), 0, ox);
    at(5, make_uint3(x1, g.y, z1), 1, oy);
    at(6, make_uint3(g.x, y1, z1), 0, ox);
    at(7, make_uint3(g.x, g.y, z1), 1, oy);
    at(8, make_uint3(g.x, g.y, g.z), 2, oz);
    at(9, make_uint3(x1, g.y, g.z), 2, oz);
    at(10, make_uint3(x1, y1, g.z), 2, oz);
    at(11, make_uint3(g.x, y1, g.z), 2, oz);
    const uint nt = c_mc.nverts[c] / 3u;
    const uint base = tscan[voxel];
    for (uint t = 0; t < nt; ++t) {
        const uint index = base + t;
        const uint a = min(ev[c_mc.tri[c][3 * t + 0]], k.maxv - 1u);
        const uint b = min(ev[c_mc.tri[c][3 * t + 1]], k.maxv - 1u);
        const uint d = min(ev[c_mc.tri[c][3 * t + 2]], k.maxv - 1u);
        if (index < k.ntris) tris[index] = make_uint3(b, d, a);
    }
}

// incidence: count, then fill (rx_mc_kernel.cu:594-626 scatters normals; here the triangle ids are scattered)
__global__ __launch_bounds__(kTpb) void k_mc_incidence_count(const uint3* __restrict__ tris, uint nt, uint nv, uint* __restrict__ cnt) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nt) return;
    const uint3 t = tris[i];
    if (t.x < nv) atomicAdd(&cnt[t.x], 1u);
    if (t.y < nv) atomicAdd(&cnt[t.y], 1u);
    if (t.z < nv) atomicAdd(&cnt[t.z], 1u);
}
__global__ __launch_bounds__(kTpb) void k_mc_incidence_fill(const uint3* __restrict__ tris, uint nt, uint nv,
                                                            const uint* __restrict__ start, uint* __restrict__ fill,
                                                            uint* __restrict__ inc) {
    const uint i = blockIdx.x * kTpb + threadIdx.x;
    if (i >= nt) return;
    const uint3 t = tris[i];
    if (t.x < nv) inc[start[t.x] + atomicAdd(&fill[t.x], 1u)] = i;
    if (t.y < nv) inc[start[t.y] + atomicAdd(&fill[t.y], 1u)] = i;
    if (t.z < nv) inc[start[t.z] + atomicAdd(&fill[t.z], 1u)] = i;
}

bool read_totals(uint* dev, uint* host, hipStream_t s) {
    return RX_HIP(hipMemcpyAsync(host, dev, 3 * sizeof(uint), hipMemcpyDeviceToHost, s)) && RX_HIP(hipStreamSynchronize(s));
}

bool mc_check(const char* fn, bool ok, const char* what) {
    if (!ok) rx_set_error(2, "%s: %s", fn, what);
    return ok;
}

}  // namespace

extern "C" {

// Tables are compiled into the library's constant memory (mc_tables.h), so there is nothing to set up.
bool CuMCInit(void) { return true; }
void CuInitMCTable(void) {}
void CuCleanMCTable(void) {}

void CuMCCalTriNum(float* dVolume, uint* dVoxBit, uint* dVoxVNum, uint* dVoxVNumScan, uint* dVoxTNum, uint* dVoxTNumScan,
                   uint* dVoxOcc, uint* dVoxOccScan, uint* dCompactedVox, uint3 grid_size, uint num_voxels, float3 grid_width,
                   float threshold, uint& num_active_voxels, uint& nvrts, uint& ntris) {
    const char* fn = "CuMCCalTriNum";
    num_active_voxels = 0;
    nvrts = 0;
    ntris = 0;
    if (num_voxels == 0) return;
    if (!mc_check(fn, dVolume && dVoxBit && dVoxVNum && dVoxVNumScan && dVoxTNum && dVoxTNumScan && dVoxOcc && dVoxOccScan &&
                          dCompactedVox, "null array") ||
        !mc_check(fn, (uint64_t)grid_size.x * grid_size.y * grid_size.z == num_voxels, "num_voxels != grid_size product"))
        return;
    hipStream_t s = rx_stream();
    uint* tot = (uint*)rx_scratch(4, 4 * sizeof(uint));
    if (!tot) return;
    uint h[3];
    k_mc_classify<<<rx_blocks(num_voxels, kTpb), kTpb, 0, s>>>(dVolume, grid_size, num_voxels, threshold, dVoxBit, dVoxVNum,
                                                                  dVoxTNum, dVoxOcc);
    rx_scan_u32(dVoxOcc, dVoxOccScan, num_voxels, false, s);
    k_scan_totals<<<1, 64, 0, s>>>(dVoxOcc, dVoxOccScan, nullptr, nullptr, nullptr, nullptr, num_voxels, tot);
    if (!read_totals(tot, h, s)) return;
    num_active_voxels = h[0];
    if (!num_active_voxels) return;  // rx_mc.cu:181-184
    k_compact_index<<<rx_blocks(num_voxels, kTpb), kTpb, 0, s>>>(dCompactedVox, dVoxOcc, dVoxOccScan, num_voxels);
    rx_scan_u32(dVoxVNum, dVoxVNumScan, num_voxels, false, s);
    rx_scan_u32(dVoxTNum, dVoxTNumScan, num_voxels, false, s);
    k_scan_totals<<<1, 64, 0, s>>>(dVoxVNum, dVoxVNumScan, dVoxTNum, dVoxTNumScan, nullptr, nullptr, num_voxels, tot);
    if (!read_totals(tot, h, s)) return;
    nvrts = h[0];
    ntris = h[1];
}

void CuMCCalEdgeVrts(float* dVolume, float* dEdgeVrts, float* dCompactedEdgeVrts, uint* dEdgeOcc, uint* dEdgeOccScan,
                     uint3 edge_size[3], uint num_edge[4], uint3 grid_size, uint num_voxels, float3 grid_width, float3 grid_min,
                     float threshold, uint& nvrts) {
    const char* fn = "CuMCCalEdgeVrts";
    nvrts = 0;
    if (!mc_check(fn, edge_size && num_edge, "null edge_size / num_edge")) return;
    const uint total = num_edge[3];
    if (total == 0) return;
    if (!mc_check(fn, dVolume && dEdgeVrts && dCompactedEdgeVrts && dEdgeOcc && dEdgeOccScan, "null array") ||
        !mc_check(fn, (uint64_t)num_edge[0] + num_edge[1] + num_edge[2] == total, "num_edge[3] != num_edge[0..2] sum"))
        return;
    for (int d = 0; d < 3; ++d)
        if (!mc_check(fn, (uint64_t)edge_size[d].x * edge_size[d].y * edge_size[d].z == num_edge[d], "edge_size / num_edge mismatch"))
            return;
    hipStream_t s = rx_stream();
    EdgeK k;
    for (int d = 0; d < 3; ++d) k.size[d] = edge_size[d];
    k.first[0] = 0;
    k.first[1] = num_edge[0];
    k.first[2] = num_edge[0] + num_edge[1];
    k.total = total;
    k.n = grid_size;
    k.h = grid_width;
    k.lo = grid_min;
    k.thr = threshold;
    k_mc_edges<<<rx_blocks(total, kTpb), kTpb, 0, s>>>(dVolume, k, (float4*)dEdgeVrts, dEdgeOcc);
    rx_scan_u32(dEdgeOcc, dEdgeOccScan, total, false, s);
    uint* tot = (uint*)rx_scratch(4, 4 * sizeof(uint));
    if (!tot) return;
    uint h[3];
    k_scan_totals<<<1, 64, 0, s>>>(dEdgeOcc, dEdgeOccScan, nullptr, nullptr, nullptr, nullptr, total, tot);
    if (!read_totals(tot, h, s)) return;
    nvrts = h[0];
    if (!nvrts) return;
    k_compact_f4<<<rx_blocks(total, kTpb), kTpb, 0, s>>>((float4*)dCompactedEdgeVrts, dEdgeOcc, dEdgeOccScan, (const float4*)dEdgeVrts,
                                                          total);
    RX_HIP(hipStreamSynchronize(s));
}

void CuMCCalTri(uint* dTris, uint* dVoxBit, uint* dVoxTNumScan, uint* dCompactedVox, uint* dEdgeOccScan, uint3 edge_size[3],
                uint num_edge[4], uint3 grid_size, uint num_voxels, float3 grid_width, float threshold, uint num_active_voxels,
                uint nvrts, uint ntris) {
    const char* fn = "CuMCCalTri";
    if (num_active_voxels == 0 || ntris == 0) return;
    if (!mc_check(fn, dTris && dVoxBit && dVoxTNumScan && dCompactedVox && dEdgeOccScan && edge_size && num_edge, "null array") ||
        !mc_check(fn, nvrts > 0, "no vertices") || !mc_check(fn, num_active_voxels <= num_voxels, "more active voxels than voxels"))
        return;
    hipStream_t s = rx_stream();
    TriK k;
    for (int d = 0; d < 3; ++d) k.esize[d] = edge_size[d];
    k.ecount = make_uint3(num_edge[0], num_edge[1], num_edge[2]);
    k.n = grid_size;
    k.active = num_active_voxels;
    k.maxv = nvrts;
    k.ntris = ntris;
    k_mc_triangles<<<rx_blocks(num_active_voxels, kTpb), kTpb, 0, s>>>((uint3*)dTris, dVoxTNumScan, dEdgeOccScan, dVoxBit,
                                                                        dCompactedVox, k);
    RX_HIP(hipStreamSynchronize(s));
}

void CuMCCalNrm(float* dNrms, uint* dTris, float* dCompactedEdgeVrts, uint nvrts, uint ntris) {
    const char* fn = "CuMCCalNrm";
    if (nvrts == 0) return;
    if (!mc_check(fn, dNrms && dCompactedEdgeVrts && (dTris || ntris == 0), "null array")) return;
    hipStream_t s = rx_stream();
    // slot 2: counts then starts (nvrts + 1) and fill cursors (nvrts); slot 3: the incidence list (3 ntris)
    uint* cnt = (uint*)rx_scratch(2, sizeof(uint) * (2 * (size_t)nvrts + 1));
    uint* fill = cnt + nvrts + 1;
    uint* inc = (uint*)rx_scratch(3, sizeof(uint) * (3 * (size_t)ntris + 1));
    if (!cnt || !inc) return;
    rx_fill<uint>(cnt, 0u, 2 * (size_t)nvrts + 1, s);
    if (ntris) k_mc_incidence_count<<<rx_blocks(ntris, kTpb), kTpb, 0, s>>>((const uint3*)dTris, ntris, nvrts, cnt);
    rx_scan_u32(cnt, cnt, nvrts, true, s);  // cnt[v] = start of v, cnt[nvrts] = total
    if (ntris) k_mc_incidence_fill<<<rx_blocks(ntris, kTpb), kTpb, 0, s>>>((const uint3*)dTris, ntris, nvrts, cnt, fill, inc);
    k_mc_vertex_normals<<<rx_blocks(nvrts, kTpb), kTpb, 0, s>>>((const float4*)dCompactedEdgeVrts, (const uint3*)dTris, cnt, inc,
                                                                  dNrms, nvrts);
    RX_HIP(hipStreamSynchronize(s));
}

}  // extern "C"

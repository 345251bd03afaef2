// rx_common.h -- device-side constants, grid helpers, error reporting and launch helpers shared by the HIP
// translation units of the particle path (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/rx_cu_common.h"

// --------------------------------------------------------------------------------------------- errors
// Thread-local "last error" slot behind CuGetLastErrorCode; messages also go to stderr, as the reference's
// RX_CUCHECK does (rx_cuda_utils.h:22-54).
void rx_set_error(int code, const char* fmt, ...);
bool rx_check(hipError_t e, const char* what);
#define RX_HIP(call) rx_check((call), #call)

// Library-owned state of the ABI: the last parameters handed to CuSetParameters and the stream used.
hipStream_t rx_stream();
const rxSimParams& rx_params();
bool rx_params_set();

// Persistent device scratch owned by the library (grown on demand, freed by CuClearData).
void* rx_scratch(int slot, size_t bytes);
void rx_scratch_free_all();

// --------------------------------------------------------------------------------------------- constants
// Constants every neighbour kernel needs, derived from rxSimParams on the host and passed by value (they
// land in the kernel-argument segment and are read with scalar loads).
struct SphK {
    float3 origin;
    float3 cellw;
    int gx, gy, gz;
    uint ncells;
    float h, h2;
    float mass;
    float rho0, inv_rho0;
    float gas_k, visc;
    float wpoly6, gwpoly6, gwspiky, lwvisc;
    float3 gravity;
    float3 wall_lo, wall_hi;  // box shrunk by the particle radius
    int ap;
    float ap_k, ap_n, inv_ap_wq;
    float dt;
};

SphK rx_make_consts(const rxSimParams& p);

// --------------------------------------------------------------------------------------------- float3 ops
__host__ __device__ inline float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__host__ __device__ inline float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
// (+, -, +=, and scalar * come with HIP's vector types)
__host__ __device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// --------------------------------------------------------------------------------------------- grid
// Cell coordinate of a point: floor((p - origin)/cell) in IEEE FP32 (hipcc's default division is correctly
// rounded), clamped into the grid -- the formula of rx_cu_common.cu:120-133.
__device__ inline int3 grid_pos(float3 p, const SphK& k) {
    int3 g;
    g.x = (int)floorf((p.x - k.origin.x) / k.cellw.x);
    g.y = (int)floorf((p.y - k.origin.y) / k.cellw.y);
    g.z = (int)floorf((p.z - k.origin.z) / k.cellw.z);
    g.x = min(max(g.x, 0), k.gx - 1);
    g.y = min(max(g.y, 0), k.gy - 1);
    g.z = min(max(g.z, 0), k.gz - 1);
    return g;
}
// rx_cu_common.cu:141-144: x fastest, then y, then z.
__device__ inline uint grid_hash(int3 g, const SphK& k) { return (uint)((g.z * k.gy + g.y) * k.gx + g.x); }

constexpr uint kEmpty = 0xFFFFFFFFu;
constexpr uint kDeletedHash = 2147483647u;  // hash of attr == -1 particles (rx_sph_kernel.cu:64)

// Contiguous slot range [s, e) of the cells (x0..x1, y, z) in sorted order, from cellStart/cellEnd with
// 0xFFFFFFFF marking empty cells.  Cells of one row are consecutive hashes, hence consecutive slots.
__device__ inline void row_range(const uint* __restrict__ cs, const uint* __restrict__ ce, uint h0, int nx, uint& s, uint& e) {
    s = kEmpty;
    e = 0;
    for (int t = 0; t < nx; ++t) {
        uint a = cs[h0 + t];
        if (a != kEmpty) {
            if (s == kEmpty) s = a;
            e = ce[h0 + t];
        }
    }
    if (s == kEmpty) { s = 0; e = 0; }
}

// --------------------------------------------------------------------------------------------- launch
inline unsigned rx_blocks(size_t n, unsigned tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// Blocks are dealt round-robin to the 8 XCDs of the MI355X; this remap gives each XCD a contiguous run of
// blocks (hence of sorted particles) so that neighbour re-reads hit the XCD's own L2.
__device__ inline uint xcd_block(uint b, uint nb) {
    uint per = nb >> 3, rem = nb & 7u;
    uint x = b & 7u, k = b >> 3;
    // blocks of XCD x get the logical range starting at x * per + min(x, rem)
    return x * per + min(x, rem) + k;
}
// Thread index of a one-dimensional launch under the remap above.  For passes over sorted particles, whose work
// per slot is even; a lattice pass keeps the round-robin order, as its empty regions would idle whole XCDs.
__device__ inline uint xcd_thread() { return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; }

// Generic grid-stride fill used instead of hipMemsetAsync (graph-safe).
template <class T> __global__ void k_fill(T* __restrict__ p, T v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
template <class T> inline void rx_fill(T* p, T v, size_t n, hipStream_t s) {
    if (n == 0) return;
    unsigned b = (unsigned)std::min<size_t>(rx_blocks(n, 256), 8192);
    k_fill<T><<<b, 256, 0, s>>>(p, v, n);
}

// Device-wide exclusive scan of uint (n elements); out may alias in; also writes the total to out[n] when
// write_total is set (out must then hold n + 1 elements).
// ws: optional caller-owned workspace of rx_scan_ws_bytes(n) bytes (required inside graph capture, where
// the library scratch must not be allocated); null uses library scratch.
void rx_scan_u32(const uint* in, uint* out, size_t n, bool write_total, hipStream_t s, void* ws = nullptr);
size_t rx_scan_ws_bytes(size_t n);
void rx_scan_f32(const float* in, float* out, size_t n, hipStream_t s);

// Stable LSD radix sort (8-bit digits) of n (key, value) pairs by the low nbits bits of the key: reads (k0, v0)
// (v0 null: the values are the positions 0 .. n-1), ping-pongs through (k1, v1) and (k2, v2) and leaves the
// sorted pairs in (k1, v1).  The first pass writes (k1, v1) when rx_radix_passes(nbits) is odd and (k2, v2)
// when it is even; the other pair may be the input pair itself (it is written only after the first pass has
// read the input).  ws: rx_radix_ws_bytes(n) bytes of device workspace (no allocation happens here, so the
// launches can be captured into a graph).
int rx_radix_passes(int nbits);
size_t rx_radix_ws_bytes(size_t n);
void rx_radix_sort_pairs(const uint* k0, const uint* v0, uint* k1, uint* v1, uint* k2, uint* v2, size_t n, int nbits, void* ws,
                         hipStream_t s);
inline int rx_bits_for(uint max_key) { return max_key ? 32 - __builtin_clz(max_key) : 1; }

// rx_cu_funcs.hip -- device/memory/scan/sort/grid part of the extern "C" particle ABI (rx_cu_funcs.cuh:24-66).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "rx_common.h"
#include "../../include/rx_cu_funcs.h"

// ============================================================================================ errors
namespace {
struct ErrSlot {
    int code = 0;
    char msg[512] = {0};
};
thread_local ErrSlot g_err;

rxSimParams g_params;
bool g_params_set = false;
hipStream_t g_stream = nullptr;
std::mutex g_mu;

struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
};
Scratch g_scratch[16];
}  // namespace

void rx_set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err.msg, sizeof(g_err.msg), fmt, ap);
    va_end(ap);
    g_err.code = code;
    fprintf(stderr, "[rx_cu] error %d: %s\n", code, g_err.msg);
}

bool rx_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    rx_set_error((int)e, "%s failed: %s", what, hipGetErrorString(e));
    return false;
}

hipStream_t rx_stream() {
    if (!g_stream) {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_stream) RX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    }
    return g_stream;
}
const rxSimParams& rx_params() { return g_params; }
bool rx_params_set() { return g_params_set; }

void* rx_scratch(int slot, size_t bytes) {
    Scratch& s = g_scratch[slot];
    if (s.bytes < bytes) {
        if (s.p) {
            RX_HIP(hipStreamSynchronize(rx_stream()));
            RX_HIP(hipFree(s.p));
        }
        size_t b = bytes + bytes / 4 + 256;
        s.p = nullptr;
        s.bytes = 0;
        if (RX_HIP(hipMalloc(&s.p, b))) s.bytes = b;
    }
    return s.p;
}
void rx_scratch_free_all() {
    if (g_stream) RX_HIP(hipStreamSynchronize(g_stream));
    for (auto& s : g_scratch) {
        if (s.p) RX_HIP(hipFree(s.p));
        s.p = nullptr;
        s.bytes = 0;
    }
}

SphK rx_make_consts(const rxSimParams& p) {
    SphK k;
    k.origin = p.WorldOrigin;
    k.cellw = p.CellWidth;
    k.gx = (int)p.GridSize.x;
    k.gy = (int)p.GridSize.y;
    k.gz = (int)p.GridSize.z;
    k.ncells = p.NumCells;
    k.h = p.EffectiveRadius;
    k.h2 = p.EffectiveRadius * p.EffectiveRadius;
    k.mass = p.Mass;
    k.rho0 = p.Density;
    k.inv_rho0 = 1.0f / p.Density;
    k.gas_k = p.GasStiffness;
    k.visc = p.Viscosity;
    k.wpoly6 = p.Wpoly6;
    k.gwpoly6 = p.GWpoly6;
    k.gwspiky = p.GWspiky;
    k.lwvisc = p.LWvisc;
    k.gravity = p.Gravity;
    k.wall_lo = f3(p.BoundaryMin.x + p.ParticleRadius, p.BoundaryMin.y + p.ParticleRadius, p.BoundaryMin.z + p.ParticleRadius);
    k.wall_hi = f3(p.BoundaryMax.x - p.ParticleRadius, p.BoundaryMax.y - p.ParticleRadius, p.BoundaryMax.z - p.ParticleRadius);
    k.ap = p.AP;
    k.ap_k = p.AP_K;
    k.ap_n = p.AP_N;
    k.inv_ap_wq = (p.AP_WQ != 0.0f) ? 1.0f / p.AP_WQ : 0.0f;
    k.dt = p.Dt;
    return k;
}

// ============================================================================================ scan
namespace {
constexpr int kScanTpb = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanTpb * kScanItems;

// Block-wide exclusive scan of one value per thread: wave64 shuffles, then the 4 wave totals through LDS.
template <class T> __device__ T block_exclusive(T v, T* wsum, T& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    T pre = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kScanTpb / 64; ++w) {
        T t = wsum[w];
        if (w < wid) pre += t;
        total += t;
    }
    __syncthreads();
    return pre + x - v;
}

template <class T> __global__ __launch_bounds__(kScanTpb) void k_scan_reduce(const T* __restrict__ in, T* __restrict__ sums, size_t n) {
    __shared__ T wsum[kScanTpb / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    T tot;
    block_exclusive(s, wsum, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T> __global__ __launch_bounds__(kScanTpb) void k_scan_apply(const T* __restrict__ in, T* __restrict__ out, const T* __restrict__ offs, size_t n) {
    __shared__ T wsum[kScanTpb / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : (T)0;
        s += v[k];
    }
    T tot;
    T pre = block_exclusive(s, wsum, tot) + (offs ? offs[blockIdx.x] : (T)0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
}

// total = exclusive[n-1] + last input element (last_in points at that element, saved before the scan)
template <class T> __global__ void k_write_total(const T* __restrict__ last_in, const T* __restrict__ out, T* total_dst, size_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total_dst = (n ? out[n - 1] + *last_in : (T)0);
}

// Recursive three-phase scan: per-tile sums, scan of the sums, rescan of the tiles with their offsets.
// The tile sums of every recursion level live in ws (rx_scan_ws_bytes), or in library scratch slots 8..
// when ws is null.
template <class T> void scan_impl(const T* in, T* out, size_t n, hipStream_t s, int level, char* ws) {
    if (n == 0) return;
    size_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        k_scan_apply<T><<<1, kScanTpb, 0, s>>>(in, out, nullptr, n);
        return;
    }
    T* sums = ws ? (T*)ws : (T*)rx_scratch(8 + level, sizeof(T) * tiles);
    if (!sums) return;  // the failed allocation is recorded as the last error
    char* next = ws ? ws + ((sizeof(T) * tiles + 255) & ~(size_t)255) : nullptr;
    k_scan_reduce<T><<<(unsigned)tiles, kScanTpb, 0, s>>>(in, sums, n);
    scan_impl<T>(sums, sums, tiles, s, level + 1, next);
    k_scan_apply<T><<<(unsigned)tiles, kScanTpb, 0, s>>>(in, out, sums, n);
}
}  // namespace

size_t rx_scan_ws_bytes(size_t n) {
    size_t b = 256;  // the saved last element
    while (n > (size_t)kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        b += ((sizeof(uint) * n + 255) & ~(size_t)255);
    }
    return b + 256;
}

__global__ void k_copy_last(const uint* __restrict__ in, uint* __restrict__ last, size_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *last = in[n - 1];
}

void rx_scan_u32(const uint* in, uint* out, size_t n, bool write_total, hipStream_t s, void* ws) {
    char* w = (char*)ws;
    if (!write_total) {
        scan_impl<uint>(in, out, n, s, 0, w ? w + 256 : nullptr);
        return;
    }
    // keep the last input element: the scan may overwrite it when out aliases in
    uint* last = w ? (uint*)w : (uint*)rx_scratch(7, sizeof(uint));
    if (!last) return;
    if (n) k_copy_last<<<1, 64, 0, s>>>(in, last, n);
    scan_impl<uint>(in, out, n, s, 0, w ? w + 256 : nullptr);
    k_write_total<uint><<<1, 64, 0, s>>>(last, out, out + n, n);
}
void rx_scan_f32(const float* in, float* out, size_t n, hipStream_t s) { scan_impl<float>(in, out, n, s, 0, nullptr); }

// ============================================================================================ grid kernels
namespace {
// rx_sph_kernel.cu:54-73 (calcHashD2)
__global__ void k_calc_hash(uint* __restrict__ hash, uint* __restrict__ index, const float4* __restrict__ pos,
                            const int* __restrict__ attr, int n, SphK k) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint hv = kDeletedHash;
    if (!attr || attr[i] != -1) hv = grid_hash(grid_pos(f3(pos[i]), k), k);
    hash[i] = hv;
    index[i] = (uint)i;
}

// rx_sph_kernel.cu:81-99 (calcHashB)
__global__ void k_calc_hash_b(uint* __restrict__ hash, uint* __restrict__ index, const float4* __restrict__ pos, int n, SphK k) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    hash[i] = grid_hash(grid_pos(f3(pos[i]), k), k);
    index[i] = (uint)i;
}

// ---- CuSort: keys are mapped to buckets (cells 0..ncells-1; bucket ncells = the deleted-particle marker and any
// other out-of-range key), the (bucket, index) pairs are radix-sorted (stable), and the buckets are mapped back.
__global__ void k_sort_bucket(const uint* __restrict__ key, uint n, uint ncells, uint* __restrict__ bucket, int* __restrict__ bad) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint h = key[i];
    uint b = h;
    if (h >= ncells) {
        b = ncells;
        if (h != kDeletedHash) atomicOr(bad, 1);
    }
    bucket[i] = b;
}
// (oval may be sv: each element is read before it is written)
__global__ void k_sort_unbucket(const uint* __restrict__ sb, const uint* sv, uint n, uint ncells, uint* __restrict__ okey, uint* oval) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint b = sb[i];
    okey[i] = b == ncells ? kDeletedHash : b;
    oval[i] = sv[i];
}

// rx_sph_kernel.cu:108-153 (reorderDataAndFindCellStartD): cell bounds from the sorted keys (empty cells were
// filled with 0xFFFFFFFF beforehand; marker keys are skipped) and the gather into sorted order.
__global__ void k_reorder(const uint* __restrict__ hash, const uint* __restrict__ index, uint* __restrict__ cstart,
                          uint* __restrict__ cend, const float4* __restrict__ pos, const float4* __restrict__ vel,
                          float4* __restrict__ spos, float4* __restrict__ svel, uint n, uint ncells) {
    uint i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint h = hash[i];
    uint prev = (i > 0) ? hash[i - 1] : kEmpty;
    uint next = (i + 1 < n) ? hash[i + 1] : kEmpty;
    if (h < ncells) {
        if (i == 0 || h != prev) cstart[h] = i;
        if (i + 1 == n || h != next) cend[h] = i + 1;
    }
    uint src = index[i];
    spos[i] = pos[src];
    if (vel) svel[i] = vel[src];
}
}  // namespace

// ============================================================================================ extern "C"
extern "C" {

int CuGetLastErrorCode(void) { return g_err.code; }
const char* CuGetLastErrorString(void) { return g_err.msg; }
void* CuGetStream(void) { return (void*)rx_stream(); }

void CuClearLastError(void) {
    g_err.code = 0;
    g_err.msg[0] = 0;
}

// rx_cu_funcs.cu:40-55: "-device=N" selects the GPU, otherwise device 0.
void CuInit(int argc, char** argv) {
    int id = 0;
    for (int a = 1; a < argc; ++a)
        if (argv[a] && strncmp(argv[a], "-device=", 8) == 0) id = atoi(argv[a] + 8);
    CuSetDevice(id);
}

static void* rx_noalloc_scratch();

void CuSetDevice(int id) {
    int count = 0;
    if (!RX_HIP(hipGetDeviceCount(&count))) return;
    if (id < 0 || id >= count) id = 0;
    if (!RX_HIP(hipSetDevice(id))) return;
    rx_noalloc_scratch();  // set-up of the allocation-free sort's static scratch on this device
}

void CuDeviceProp(void) {
    int n = 0;
    if (!RX_HIP(hipGetDeviceCount(&n))) return;
    for (int i = 0; i < n; ++i) {
        hipDeviceProp_t d;
        if (!RX_HIP(hipGetDeviceProperties(&d, i))) continue;
        printf("device %d: %s (%s), %d CUs, %.1f GiB, clock %d kHz, mem clock %d kHz, bus %d bit, L2 %d KiB\n", i, d.name,
               d.gcnArchName, d.multiProcessorCount, d.totalGlobalMem / 1073741824.0, d.clockRate, d.memoryClockRate,
               d.memoryBusWidth, d.l2CacheSize / 1024);
    }
}

void CuAllocateArray(void** devPtr, int size) {
    if (size < 0) {
        rx_set_error(1, "CuAllocateArray: negative size %d", size);
        *devPtr = nullptr;
        return;
    }
    RX_HIP(hipMalloc(devPtr, (size_t)size));
}
void CuSetArrayValue(void* devPtr, int val, size_t size) {
    if (!size) return;
    RX_HIP(hipMemsetAsync(devPtr, val, size, rx_stream()));
    RX_HIP(hipStreamSynchronize(rx_stream()));
}
void CuFreeArray(void* devPtr) {
    if (!devPtr) return;
    RX_HIP(hipStreamSynchronize(rx_stream()));
    RX_HIP(hipFree(devPtr));
}
void CuCopyArrayD2D(void* dDst, void* dSrc, int size) {
    if (size <= 0) return;
    RX_HIP(hipMemcpyAsync(dDst, dSrc, (size_t)size, hipMemcpyDeviceToDevice, rx_stream()));
    RX_HIP(hipStreamSynchronize(rx_stream()));
}
void CuThreadSync(void) {
    RX_HIP(hipStreamSynchronize(rx_stream()));
    RX_HIP(hipDeviceSynchronize());
}
void CuCopyArrayFromDevice(void* host, const void* device, cudaGraphicsResource** resource, int size) {
    if (resource) {
        rx_set_error(2, "CuCopyArrayFromDevice: GL buffer objects are not supported");
        return;
    }
    if (size <= 0) return;
    RX_HIP(hipMemcpyAsync(host, device, (size_t)size, hipMemcpyDeviceToHost, rx_stream()));
    RX_HIP(hipStreamSynchronize(rx_stream()));
}
void CuCopyArrayToDevice(void* device, const void* host, int offset, int size) {
    if (size <= 0) return;
    RX_HIP(hipMemcpyAsync((char*)device + offset, host, (size_t)size, hipMemcpyHostToDevice, rx_stream()));
    RX_HIP(hipStreamSynchronize(rx_stream()));
}
void CuRegisterGLBufferObject(unsigned int vbo, cudaGraphicsResource** resource) {
    if (resource) *resource = nullptr;
    rx_set_error(2, "CuRegisterGLBufferObject: GL buffer objects are not supported (vbo %u)", vbo);
}
void CuUnregisterGLBufferObject(cudaGraphicsResource* resource) {
    rx_set_error(2, "CuUnregisterGLBufferObject: GL buffer objects are not supported");
}
void* CuMapGLBufferObject(cudaGraphicsResource** resource) {
    rx_set_error(2, "CuMapGLBufferObject: GL buffer objects are not supported");
    return nullptr;
}
void CuUnmapGLBufferObject(cudaGraphicsResource* resource) {
    rx_set_error(2, "CuUnmapGLBufferObject: GL buffer objects are not supported");
}

void CuScan(unsigned int* output, unsigned int* input, unsigned int numElements) {
    rx_scan_u32(input, output, numElements, false, rx_stream());
    RX_HIP(hipStreamSynchronize(rx_stream()));
}
void CuScanf(float* output, float* input, unsigned int numElements) {
    rx_scan_f32(input, output, numElements, rx_stream());
    RX_HIP(hipStreamSynchronize(rx_stream()));
}

// Stable sort of (hash, index) pairs (replaces thrust::sort_by_key of rx_cu_funcs.cu:255-260): ties keep their
// input order.  Scratch: 3 n uints (the buckets and one ping-pong pair; the buckets and dIndex are the other
// pair) + the radix workspace.
size_t rx_sort_scratch_bytes(size_t n) { return ((sizeof(uint) * 3 * n + 255) & ~(size_t)255) + rx_radix_ws_bytes(n); }

// The launches of CuSort on stream s with every buffer supplied (scratch: rx_sort_scratch_bytes(n) bytes, 256-byte
// aligned).  Sets the device int `bad` when a key is neither a cell hash nor the deleted marker.
void rx_sort_launch(uint* dHash, uint* dIndex, uint n, uint nc, void* scratch, int* bad, hipStream_t s) {
    uint* bucket = (uint*)scratch;
    uint* kx = bucket + (size_t)n;
    uint* vx = bucket + 2 * (size_t)n;
    void* rws = (char*)scratch + ((sizeof(uint) * 3 * (size_t)n + 255) & ~(size_t)255);
    const unsigned blocks = rx_blocks(n, 256);
    const int nbits = rx_bits_for(nc);
    k_sort_bucket<<<blocks, 256, 0, s>>>(dHash, n, nc, bucket, bad);
    if (rx_radix_passes(nbits) & 1) {  // the first pass writes (kx, vx), where the result ends
        rx_radix_sort_pairs(bucket, dIndex, kx, vx, bucket, dIndex, n, nbits, rws, s);
        k_sort_unbucket<<<blocks, 256, 0, s>>>(kx, vx, n, nc, dHash, dIndex);
    } else {  // the first pass writes (kx, vx), the last one (bucket, dIndex)
        rx_radix_sort_pairs(bucket, dIndex, bucket, dIndex, kx, vx, n, nbits, rws, s);
        k_sort_unbucket<<<blocks, 256, 0, s>>>(bucket, dIndex, n, nc, dHash, dIndex);
    }
}

void CuSort(unsigned int* dHash, uint* dIndex, uint numParticles) {
    if (numParticles == 0) return;
    if (!g_params_set) {
        rx_set_error(3, "CuSort: CuSetParameters must be called first (the bucket count is NumCells)");
        return;
    }
    hipStream_t s = rx_stream();
    // keys are cell hashes of the fluid grid or of the boundary-particle grid (CuCalcHashB)
    const uint n = numParticles, nc = std::max(g_params.NumCells, g_params.NumCellsB);
    void* scratch = rx_scratch(1, rx_sort_scratch_bytes(n) + 256);
    int* bad = (int*)rx_scratch(0, 256);
    if (!scratch || !bad) return;
    rx_fill<int>(bad, 0, 1, s);
    rx_sort_launch(dHash, dIndex, n, nc, scratch, bad, s);
    int hbad = 0;
    RX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    RX_HIP(hipStreamSynchronize(s));
    if (hbad) rx_set_error(4, "CuSort: keys other than cell hashes (< %u) or 0x7FFFFFFF were sorted as 0x7FFFFFFF", nc);
}

// ---- the allocation-free sort: its scratch is a static device array (one copy per GPU, held until the process
// ends), whose address on the current device is looked up once per device.
namespace {
constexpr size_t kNoAllocMaxN = (size_t)1 << 21;
constexpr size_t kNoAllocFlag = 256;  // the out-of-range flag, then the sort's scratch
constexpr size_t kNoAllocBytes = kNoAllocFlag + ((3 * sizeof(uint) * kNoAllocMaxN + 255) & ~(size_t)255) + sizeof(uint) * 256 * 256;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices];
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on the current device (looked up on the first use per device).
static void* rx_noalloc_scratch() {
    int dev = 0;
    if (!RX_HIP(hipGetDevice(&dev))) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        rx_set_error(1, "CuSortNoAlloc: device %d is beyond the %d devices the library tracks", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            rx_set_error((int)e, "CuSortNoAlloc: hipGetSymbolAddress of the static scratch failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        if (rx_sort_scratch_bytes(kNoAllocMaxN) + kNoAllocFlag > kNoAllocBytes) {  // guards the size formula above
            rx_set_error(1, "CuSortNoAlloc: static scratch smaller than the sort needs");
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

void CuSortNoAlloc(unsigned int* dHash, uint* dIndex, uint numParticles) {
    if (numParticles == 0) return;
    if ((size_t)numParticles > kNoAllocMaxN) {
        rx_set_error(1, "CuSortNoAlloc: %u keys exceed the static scratch (at most %zu)", numParticles, kNoAllocMaxN);
        return;
    }
    if (!g_params_set) {
        rx_set_error(3, "CuSortNoAlloc: CuSetParameters must be called first (the bucket count is NumCells)");
        return;
    }
    char* base = (char*)rx_noalloc_scratch();
    if (!base) return;
    hipStream_t s = rx_stream();
    const uint nc = std::max(g_params.NumCells, g_params.NumCellsB);
    rx_fill<int>((int*)base, 0, 1, s);  // a fill kernel, not a memset (the flag starts at zero on every call)
    rx_sort_launch(dHash, dIndex, numParticles, nc, base + kNoAllocFlag, (int*)base, s);
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void CuSPHInit(int max_particles) {
    if (max_particles > 0) {
        rx_scratch(0, 256);
        rx_scratch(1, rx_sort_scratch_bytes((size_t)max_particles) + 256);
    }
}
void CuSPHClean(void) { rx_scratch_free_all(); }

void CuSetParameters(rxSimParams* hostParams) {
    if (!hostParams) {
        rx_set_error(1, "CuSetParameters: null parameters");
        return;
    }
    g_params = *hostParams;
    g_params_set = true;
}
void CuClearData(void) { rx_scratch_free_all(); }

void CuCalcHash(uint* gridParticleHash, uint* gridParticleIndex, float* pos, int* attr, int numParticles) {
    if (numParticles <= 0) return;
    if (!g_params_set) { rx_set_error(3, "CuCalcHash: CuSetParameters must be called first"); return; }
    hipStream_t s = rx_stream();
    k_calc_hash<<<rx_blocks(numParticles, 256), 256, 0, s>>>(gridParticleHash, gridParticleIndex, (const float4*)pos, attr,
                                                               numParticles, rx_make_consts(g_params));
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void CuCalcHashB(uint* dGridParticleHash, uint* dSortedIndex, float* dPos, float3 world_origin, float3 cell_width,
                 uint3 grid_size, int nprts) {
    if (nprts <= 0) return;
    SphK k{};
    k.origin = world_origin;
    k.cellw = cell_width;
    k.gx = (int)grid_size.x;
    k.gy = (int)grid_size.y;
    k.gz = (int)grid_size.z;
    hipStream_t s = rx_stream();
    k_calc_hash_b<<<rx_blocks(nprts, 256), 256, 0, s>>>(dGridParticleHash, dSortedIndex, (const float4*)dPos, nprts, k);
    RX_HIP(hipStreamSynchronize(s));
}

static void reorder_impl(rxParticleCell& cell, float* oldPos, float* oldVel) {
    uint n = cell.uNumParticles;
    hipStream_t s = rx_stream();
    rx_fill<uint>(cell.dCellStart, kEmpty, cell.uNumCells, s);
    rx_fill<uint>(cell.dCellEnd, kEmpty, cell.uNumCells, s);
    if (n) {
        k_reorder<<<rx_blocks(n, 256), 256, 0, s>>>(cell.dGridParticleHash, cell.dSortedIndex, cell.dCellStart, cell.dCellEnd,
                                                     (const float4*)oldPos, (const float4*)oldVel, cell.dSortedPos,
                                                     cell.dSortedVel, n, cell.uNumCells);
    }
    RX_HIP(hipGetLastError());
    RX_HIP(hipStreamSynchronize(s));
}

void CuReorderDataAndFindCellStart(rxParticleCell cell, float* oldPos, float* oldVel) { reorder_impl(cell, oldPos, oldVel); }
void CuReorderDataAndFindCellStartB(rxParticleCell cell, float* oldPos) { reorder_impl(cell, oldPos, nullptr); }

}  // extern "C"

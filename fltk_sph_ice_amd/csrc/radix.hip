// radix.hip -- stable least-significant-digit radix sort of (uint key, uint value) pairs.
//
// A pass sorts by one 8-bit digit.  The input is cut into at most 256 equal tiles, one 1024-thread block each
// (16 wavefronts); wavefront w of a tile owns a contiguous run of the tile and walks it in rounds of 64 (4
// rounds' loads issued together), so (wave, round, lane) is the input order.  Within a round the lanes
// holding the same digit are found with 8 ballots (one per digit bit); the lowest such lane adds their count to
// the wave's digit counter in LDS, and each lane's rank is the number of same-digit lanes below it.  A pass is
// two launches:
//   k_radix_hist     per tile and digit, the element count          -> cnt[tile * 256 + digit]
//   k_radix_scatter  each block first turns the count matrix (at most 256 x 256, L2-resident) into its tile's
//                    first output slot per digit (all counts of smaller digits plus the counts of its digit in
//                    earlier tiles), then moves every element to that slot + (count of its digit in earlier
//                    waves of the tile) + (count in earlier rounds of its wave) + its rank in the round
// Every pass is stable, so the whole sort is stable: ties keep their input order.  Work and traffic are linear
// in n whatever the key distribution (one key for all elements included).
#include "radix.h"

namespace {

constexpr int kRWaves = kRThreads / 64;
constexpr int kRBatch = 4;   // rounds whose loads are issued together
constexpr int kRDig = 1 << kRBits;
constexpr int kRSplit = kRThreads / kRDig;  // threads per digit in the offset step
constexpr int kRCntLoads = kRMaxTiles / kRSplit / 2;  // count-matrix loads in flight per thread in the offset step

// Lanes of the wavefront that are valid and hold digit d (nbits bits).
__device__ inline unsigned long long digit_match(uint d, bool valid, int nbits) {
    unsigned long long m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
        const unsigned long long v = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? v : ~v;
    }
    return m;
}

// Digit counts of this wavefront's run [wbase, wend) into wcnt[wave * ndig + digit] (zeroed by the caller).
// key_of(i) gives the key of element i.
template <class KeyOf>
__device__ inline void wave_counts(KeyOf&& key_of, size_t wbase, size_t wend, int shift, int nbits, uint* wcnt) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint mask = (1u << nbits) - 1u;
    for (size_t r0 = wbase; r0 < wend; r0 += 64 * kRBatch) {
        uint d[kRBatch];
#pragma unroll
        for (int r = 0; r < kRBatch; ++r) {
            const size_t i = r0 + (size_t)r * 64 + lane;
            d[r] = i < wend ? (key_of(i) >> shift) & mask : 0u;
        }
#pragma unroll
        for (int r = 0; r < kRBatch; ++r) {
            const size_t i = r0 + (size_t)r * 64 + lane;
            const bool valid = i < wend;
            const unsigned long long m = digit_match(d[r], valid, nbits);
            if (valid && lane == __builtin_ctzll(m)) wcnt[(w << nbits) + d[r]] += (uint)__builtin_popcountll(m);
        }
    }
}

template <class KeyOf>
__device__ inline void tile_hist(KeyOf&& key_of, size_t n, int shift, int nbits, RadixShape rs, uint* __restrict__ cnt, uint* wcnt) {
    const int ndig = 1 << nbits;
    for (int t = threadIdx.x; t < (kRWaves << nbits); t += kRThreads) wcnt[t] = 0u;
    __syncthreads();
    const size_t wbase = (size_t)blockIdx.x * rs.tile + (size_t)(threadIdx.x >> 6) * rs.wave_items;
    wave_counts(key_of, wbase, std::min(n, wbase + rs.wave_items), shift, nbits, wcnt);
    __syncthreads();
    for (int d = threadIdx.x; d < ndig; d += kRThreads) {
        uint s = 0;
        for (int w = 0; w < kRWaves; ++w) s += wcnt[(w << nbits) + d];
        cnt[(size_t)blockIdx.x * ndig + d] = s;
    }
}

__global__ __launch_bounds__(kRThreads) void k_radix_hist(const uint* __restrict__ keys, size_t n, int shift, RadixShape rs,
                                                          uint* __restrict__ cnt) {
    __shared__ uint wcnt[kRWaves * kRDig];
    tile_hist([&](size_t i) { return keys[i]; }, n, shift, kRBits, rs, cnt, wcnt);
}

// The fused-grid rebuild's first histogram: hashes the positions (written to keys for the scatter) and counts
// the first digit.
__global__ __launch_bounds__(kRThreads) void k_radix_hist_hash(const float4* __restrict__ pos, FGridK g, uint* __restrict__ keys, size_t n,
                                                               RadixShape rs, uint* __restrict__ cnt) {
    __shared__ uint wcnt[kRWaves * kRDig];
    tile_hist(
        [&](size_t i) {
            const uint h = fg_hash(fg_pos(f3(pos[i]), g), g);
            keys[i] = h;
            return h;
        },
        n, 0, kRBits, rs, cnt, wcnt);
}

// Scatter of one pass.  vals == nullptr: the values are the input positions (an identity permutation).  With
// gs != nullptr (last pass of a fused-grid rebuild) every state array is gathered into the output order too.
__global__ __launch_bounds__(kRThreads) void k_radix_scatter(const uint* __restrict__ keys, const uint* __restrict__ vals, size_t n,
                                                             int shift, RadixShape rs, const uint* __restrict__ cnt,
                                                             uint* __restrict__ okeys, uint* __restrict__ ovals, RadixGather gs) {
    constexpr int nbits = kRBits;
    __shared__ uint wcnt[kRWaves * kRDig];
    __shared__ uint part[2][kRSplit][kRDig];
    __shared__ uint base_of[kRDig];
    __shared__ uint wtot[kRDig / 64];
    for (int t = threadIdx.x; t < (kRWaves << nbits); t += kRThreads) wcnt[t] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t wbase = (size_t)blockIdx.x * rs.tile + (size_t)w * rs.wave_items;
    const size_t wend = std::min(n, wbase + rs.wave_items);
    wave_counts([&](size_t i) { return keys[i]; }, wbase, wend, shift, nbits, wcnt);
    {
        // this tile's first output slot per digit from the count matrix cnt[tile][digit]: kRSplit threads per digit
        // sum interleaved tiles (the counts of earlier tiles, and of all tiles for the digit totals)
        const int d = threadIdx.x % kRDig, q = threadIdx.x / kRDig;
        uint below = 0, all = 0;
        for (uint t0 = q; t0 < rs.ntiles; t0 += kRCntLoads * kRSplit) {  // two batches of loads per thread
            uint c[kRCntLoads];
#pragma unroll
            for (int u = 0; u < kRCntLoads; ++u) {
                const uint t = t0 + u * kRSplit;
                c[u] = t < rs.ntiles ? cnt[(size_t)t * kRDig + d] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kRCntLoads; ++u) {
                all += c[u];
                below += t0 + u * kRSplit < blockIdx.x ? c[u] : 0u;
            }
        }
        part[0][q][d] = below;
        part[1][q][d] = all;
    }
    __syncthreads();
    if (threadIdx.x < kRDig) {  // exclusive scan of the digit totals (4 wavefronts of 64 digits), plus the tile prefix
        const int d = threadIdx.x, lane = d & 63;
        uint below = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < kRSplit; ++q) {
            below += part[0][q][d];
            tot += part[1][q][d];
        }
        uint x = tot;
#pragma unroll
        for (int k = 1; k < 64; k <<= 1) {
            const uint y = __shfl_up(x, k, 64);
            if (lane >= k) x += y;
        }
        if (lane == 63) wtot[d >> 6] = x;
        base_of[d] = x - tot + below;
    }
    __syncthreads();
    // first output slot of each (wave, digit): the tile's slot plus the counts of earlier waves
    if (threadIdx.x < kRDig) {
        const int d = threadIdx.x;
        uint run = base_of[d];
        for (int ww = 0; ww < (d >> 6); ++ww) run += wtot[ww];
        for (int ww = 0; ww < kRWaves; ++ww) {
            const uint c = wcnt[(ww << nbits) + d];
            wcnt[(ww << nbits) + d] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint mask = (1u << nbits) - 1u;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint* mine = wcnt + (w << nbits);
    for (size_t r0 = wbase; r0 < wend; r0 += 64 * kRBatch) {
        uint kb[kRBatch], vb[kRBatch];  // the loads of kRBatch rounds are issued before the first round is placed
#pragma unroll
        for (int u = 0; u < kRBatch; ++u) {
            const size_t i = r0 + (size_t)u * 64 + lane;
            kb[u] = i < wend ? keys[i] : 0u;
            vb[u] = i < wend ? (vals ? vals[i] : (uint)i) : 0u;
        }
#pragma unroll
        for (int u = 0; u < kRBatch; ++u) {
            const bool valid = r0 + (size_t)u * 64 + lane < wend;
            const uint key = kb[u], val = vb[u];
            const uint d = (key >> shift) & mask;
            const unsigned long long m = digit_match(d, valid, nbits);
            // every lane reads the running slot before the lowest lane of its digit advances it (a wavefront's
            // LDS accesses complete in program order)
            const uint base = mine[d];
            if (valid) {
                const uint pos = base + (uint)__builtin_popcountll(m & below);
                okeys[pos] = key;
                ovals[pos] = val;
                if (gs.n4) {
#pragma unroll
                    for (int a = 0; a < kRGatherMax; ++a)
                        if (a < gs.n4) gs.dst4[a][pos] = gs.src4[a][val];
                    gs.dstu[pos] = gs.srcu[val];
                    gs.zero[pos] = 0.0f;
                }
                if (lane == __builtin_ctzll(m)) mine[d] = base + (uint)__builtin_popcountll(m);
            }
        }
    }
}

// off[c] = first position of a key >= c in the sorted keys (c = 0 .. ncells; off[ncells] = n).
__global__ void k_cell_offsets(const uint* __restrict__ sorted, uint n, uint ncells, uint* __restrict__ off) {
    const uint c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    uint lo = 0, hi = n;
    while (lo < hi) {
        const uint mid = (lo + hi) >> 1;
        if (sorted[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    off[c] = lo;
}

int passes_for(int nbits) { return (std::max(1, std::min(nbits, 32)) + kRBits - 1) / kRBits; }

}  // namespace

RadixShape rx_radix_shape(size_t n) {
    RadixShape r;
    const size_t per_tile = (n + kRMaxTiles - 1) / kRMaxTiles;
    r.wave_items = (uint)std::max<size_t>(64, ((per_tile + kRWaves - 1) / kRWaves + 63) / 64 * 64);
    r.tile = r.wave_items * kRWaves;
    r.ntiles = (uint)((n + r.tile - 1) / r.tile);
    return r;
}

int rx_radix_passes(int nbits) { return passes_for(nbits); }

size_t rx_radix_ws_bytes(size_t n) { return sizeof(uint) * (size_t)kRDig * kRMaxTiles; }

void rx_radix_sort_pairs(const uint* k0, const uint* v0, uint* k1, uint* v1, uint* k2, uint* v2, size_t n, int nbits, void* ws,
                         hipStream_t s) {
    if (n == 0) return;
    const RadixShape rs = rx_radix_shape(n);
    const int passes = passes_for(nbits);
    uint* cnt = (uint*)ws;
    const uint *ik = k0, *iv = v0;
    for (int p = 0; p < passes; ++p) {
        const bool to1 = ((passes - 1 - p) & 1) == 0;  // the last pass writes (k1, v1)
        uint* ok = to1 ? k1 : k2;
        uint* ov = to1 ? v1 : v2;
        k_radix_hist<<<rs.ntiles, kRThreads, 0, s>>>(ik, n, p * kRBits, rs, cnt);
        k_radix_scatter<<<rs.ntiles, kRThreads, 0, s>>>(ik, iv, n, p * kRBits, rs, cnt, ok, ov, RadixGather{});
        ik = ok;
        iv = ov;
    }
}

void rx_grid_sort_gather(const float4* pos, const FGridK& g, uint n, uint* keys, uint* hash, uint* perm, uint* k2, uint* v2, uint* off,
                         void* ws, const RadixGather& gs, hipStream_t s) {
    if (n == 0) return;
    const RadixShape rs = rx_radix_shape(n);
    const int passes = passes_for(g.ncells > 1 ? 32 - __builtin_clz(g.ncells - 1) : 1);
    uint* cnt = (uint*)ws;
    const uint *ik = keys, *iv = nullptr;
    for (int p = 0; p < passes; ++p) {
        const bool to1 = ((passes - 1 - p) & 1) == 0;  // the last pass writes (hash, perm)
        uint* ok = to1 ? hash : k2;
        uint* ov = to1 ? perm : v2;
        if (p == 0) k_radix_hist_hash<<<rs.ntiles, kRThreads, 0, s>>>(pos, g, keys, n, rs, cnt);
        else k_radix_hist<<<rs.ntiles, kRThreads, 0, s>>>(ik, n, p * kRBits, rs, cnt);
        // the values of the first pass are the current slots; every later pass carries them along, and the last
        // pass gathers the state through them
        k_radix_scatter<<<rs.ntiles, kRThreads, 0, s>>>(ik, iv, n, p * kRBits, rs, cnt, ok, ov, p == passes - 1 ? gs : RadixGather{});
        ik = ok;
        iv = ov;
    }
    k_cell_offsets<<<rx_blocks((size_t)g.ncells + 1, 256), 256, 0, s>>>(hash, n, g.ncells, off);
}

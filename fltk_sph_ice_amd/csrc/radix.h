// radix.h -- the stable radix sort's launch shape and the fused-grid rebuild built on it.
#pragma once

#include "fused_grid.h"

constexpr int kRBits = 8;                     // digit of a pass
constexpr int kRThreads = 1024;               // threads per tile (16 wavefronts)
constexpr unsigned kRMaxTiles = 256;          // tiles per pass (the count matrix is at most 256 x 256)
constexpr int kRGatherMax = 4;

struct RadixShape {
    unsigned ntiles;      // tiles (= blocks)
    unsigned wave_items;  // elements per wavefront (a multiple of 64)
    unsigned tile;        // elements per tile
};
RadixShape rx_radix_shape(size_t n);

// State arrays the last pass of a grid rebuild gathers into the new order: dst4[a][new] = src4[a][old] for
// a < n4, dstu[new] = srcu[old], zero[new] = 0.  n4 = 0: no gather.
struct RadixGather {
    int n4 = 0;
    const float4* src4[kRGatherMax];
    float4* dst4[kRGatherMax];
    const uint* srcu;
    uint* dstu;
    float* zero;
};

// Fused-grid rebuild: hashes pos (n slots, current order) into keys, sorts (cell, slot) pairs stably by cell
// into (hash, perm) -- hash = cell of each new slot, perm = new slot -> current slot -- through the ping-pong pair
// (k2, v2), gathers the state given in gs during the last pass, and writes off[c] = first new slot of cell c
// for c = 0 .. g.ncells (off[g.ncells] = n).  ws: rx_radix_ws_bytes(n) bytes.
void rx_grid_sort_gather(const float4* pos, const FGridK& g, unsigned n, unsigned* keys, unsigned* hash, unsigned* perm, unsigned* k2,
                         unsigned* v2, unsigned* off, void* ws, const RadixGather& gs, hipStream_t s);

// Packed per-tile counts of the particle tile lists, shared by the
// particle kernels and the tile-restricted field preview.
#pragma once

#include <hip/hip_runtime.h>

namespace aprl_gpu {

// Per-tile counts packed in one 64-bit word, so that one exclusive scan
// over the tiles gives a tile's first chunk, its slot among the occupied
// tiles (its fixed-point partial box) and its slot among the active tiles:
// chunks in bits 0-21, occupied (0 or 1) in bits 22-42, active in 43-63.
// The fields cannot carry into each other while the totals fit them
// (checked on the host: kMaxTiles tiles, kMaxChunks chunks).
constexpr int kChunkFieldBits = 22, kTileFieldBits = 21;
constexpr unsigned long long kMaxChunks = (1ull << kChunkFieldBits) - 1;
constexpr unsigned long long kMaxTiles = (1ull << kTileFieldBits) - 1;
__host__ __device__ constexpr unsigned long long packTileCounts(unsigned chunks, unsigned occupied, unsigned active) {
  return static_cast<unsigned long long>(chunks) | (static_cast<unsigned long long>(occupied) << kChunkFieldBits) |
         (static_cast<unsigned long long>(active) << (kChunkFieldBits + kTileFieldBits));
}
__host__ __device__ inline int chunkField(unsigned long long v) {
  return static_cast<int>(v & kMaxChunks);
}
__host__ __device__ inline int occupiedField(unsigned long long v) {
  return static_cast<int>((v >> kChunkFieldBits) & kMaxTiles);
}
__host__ __device__ inline int activeField(unsigned long long v) {
  return static_cast<int>(v >> (kChunkFieldBits + kTileFieldBits));
}

}  // namespace aprl_gpu

// Stable LSD radix sort of (tile key, particle index) pairs over 8-bit
// digits, and the device-wide exclusive scans that it and the chunk lists
// use. No library sort or scan runs on the step's path.
//
// One sort pass: kDigitHistogram counts the digits of `items` consecutive
// keys per workgroup into a digit-major table hist[d * blocks + b]; the
// exclusive scan of that table gives, for workgroup b and digit d, the
// output position of the workgroup's first item of digit d (every item of
// a smaller digit, then digit d's items of earlier workgroups);
// kDigitScatter adds each item's rank among the workgroup's earlier items
// of the same digit: inside a wave from eight ballots (the lanes of equal
// digit) and a popcount of the lower lanes, across the waves and rounds of
// the workgroup from per-digit running offsets in LDS. Ranks follow the
// input order, so every pass is stable and the permutation does not depend
// on timing. Keys of `bits` bits take ceil(bits / 8) passes.
//
// Scans: reduce-then-scan over tiles of kScanTile values: per-tile sums,
// then every tile rescanned with its offset, which the tile's workgroup
// sums from the tile sums before it (above 1024 tiles one workgroup scans
// the sums in a launch of its own; a single launch when the input fits one
// tile). Inside a
// workgroup, wavefront inclusive scans by lane shuffles and one LDS step
// over the four waves.
#include <algorithm>

#include "particles.hpp"

namespace aprl_gpu {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;
// Up to this many tiles, a tile's offset is summed by its own workgroup
// (at most 4 loads per lane) instead of by a separate one-workgroup scan.
constexpr std::size_t kSelfOffsetTiles = 4 * kScanThreads;

template <class T>
__device__ __forceinline__ T waveInclusiveScan(T v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const T t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan of one value per lane over the workgroup; total = the sum
// of all lanes. ws holds kScanWaves + 1 values.
template <class T>
__device__ __forceinline__ T blockExclusiveScan(T v, T* ws, T& total) {
  const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
  const T inc = waveInclusiveScan(v, lane);
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  if (t == 0) {
    T run = T(0);
#pragma unroll
    for (int q = 0; q < kScanWaves; ++q) {
      const T c = ws[q];
      ws[q] = run;
      run += c;
    }
    ws[kScanWaves] = run;
  }
  __syncthreads();
  const T ex = inc - v + ws[w];
  total = ws[kScanWaves];
  __syncthreads();
  return ex;
}

// Sum of every kScanTile-value tile of the input.
template <class T>
__global__ void __launch_bounds__(kScanThreads) kScanTileSums(const T* in, std::size_t n, T* sums) {
  __shared__ T ws[kScanWaves + 1];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile;
  T s = T(0);
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    const std::size_t q = base + static_cast<std::size_t>(r) * kScanThreads + threadIdx.x;
    if (q < n) s += in[q];
  }
  T total;
  (void)blockExclusiveScan(s, ws, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// In-place exclusive scan of the tile sums by one workgroup.
template <class T>
__global__ void __launch_bounds__(kScanThreads) kScanSums(T* sums, int count) {
  __shared__ T ws[kScanWaves + 1];
  T carry = T(0);
  for (int b0 = 0; b0 < count; b0 += kScanThreads) {
    const int i = b0 + static_cast<int>(threadIdx.x);
    const T v = i < count ? sums[i] : T(0);
    T total;
    const T ex = blockExclusiveScan(v, ws, total);
    if (i < count) sums[i] = ex + carry;
    carry += total;
  }
}

// Exclusive scan of one tile plus its offset: offsets[tile] (scanned tile
// sums), or with sums the sum of the tile sums before this tile, formed by
// the workgroup itself (exact in any order: integers), or 0. The tile passes
// through LDS so that global loads and stores stay coalesced while every
// lane scans kScanPer consecutive values.
template <class T>
__global__ void __launch_bounds__(kScanThreads) kScanTiles(const T* in, std::size_t n, const T* offsets,
                                                           const T* sums, T* out) {
  __shared__ T tile[kScanTile];
  __shared__ T ws[kScanWaves + 1];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile;
  const int t = static_cast<int>(threadIdx.x);
  T before = T(0);
  if (sums) {
    T part = T(0);
    for (int q = t; q < static_cast<int>(blockIdx.x); q += kScanThreads) part += sums[q];
    (void)blockExclusiveScan(part, ws, before);
  }
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    const int q = r * kScanThreads + t;
    tile[q] = base + q < n ? in[base + q] : T(0);
  }
  __syncthreads();
  T v[kScanPer];
  T s = T(0);
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    v[r] = tile[t * kScanPer + r];
    s += v[r];
  }
  T total;
  T ex = blockExclusiveScan(s, ws, total);
  if (offsets) ex += offsets[blockIdx.x];
  ex += before;
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    tile[t * kScanPer + r] = ex;
    ex += v[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    const int q = r * kScanThreads + t;
    if (base + q < n) out[base + q] = tile[q];
  }
}

template <class T>
std::size_t scanTiles(std::size_t n) {
  return (n + kScanTile - 1) / kScanTile;
}

template <class T>
void exclusiveScanImpl(const T* in, T* out, std::size_t n, void* temp, std::size_t bytes, hipStream_t st) {
  if (!n) return;
  const std::size_t tiles = scanTiles<T>(n);
  if (tiles == 1) {
    kScanTiles<T><<<1, kScanThreads, 0, st>>>(in, n, nullptr, nullptr, out);
    launchCheck("scan");
    return;
  }
  if (!temp || bytes < tiles * sizeof(T)) throw std::logic_error("Scan scratch too small");
  T* sums = static_cast<T*>(temp);
  kScanTileSums<T><<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(in, n, sums);
  if (tiles <= kSelfOffsetTiles) {
    // each tile sums the tile sums before it: no separate scan of the sums
    kScanTiles<T><<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(in, n, nullptr, sums, out);
  } else {
    kScanSums<T><<<1, kScanThreads, 0, st>>>(sums, static_cast<int>(tiles));
    kScanTiles<T><<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(in, n, sums, nullptr, out);
  }
  launchCheck("scan");
}

constexpr int kSortThreads = 256;  // one item per lane and round
constexpr int kSortWaves = kSortThreads / 64;

// Digit histogram of `items` consecutive keys per workgroup, digit-major:
// hist[digit * blocks + block].
__global__ void __launch_bounds__(kSortThreads) kDigitHistogram(const unsigned* key, int n, int shift, int items,
                                                                unsigned* hist, int blocks) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const int base = blockIdx.x * items;
  const int end = min(n, base + items);
  for (int p = base + static_cast<int>(threadIdx.x); p < end; p += kSortThreads)
    atomicAdd(&h[(key[p] >> shift) & 255u], 1u);
  __syncthreads();
  hist[static_cast<std::size_t>(threadIdx.x) * blocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one digit pass. Workgroup b owns input items
// [b * items, (b + 1) * items), taken in rounds of kSortThreads in input
// order; start = the scanned digit-major table.
__global__ void __launch_bounds__(kSortThreads) kDigitScatter(const unsigned* keyIn, const unsigned* valIn,
                                                              unsigned* keyOut, unsigned* valOut, int n, int shift,
                                                              int items, const unsigned* startTable, int blocks) {
  __shared__ unsigned waveCount[kSortWaves][256];
  __shared__ unsigned running[256];
  __shared__ unsigned start[256];
  const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
  start[t] = startTable[static_cast<std::size_t>(t) * blocks + blockIdx.x];
  running[t] = 0u;
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int r = 0; r < items; r += kSortThreads) {
    const int p = static_cast<int>(blockIdx.x) * items + r + t;
    const bool valid = r + t < items && p < n;
    const unsigned k = valid ? keyIn[p] : 0u;
    const unsigned v = valid ? valIn[p] : 0u;
    const unsigned digit = (k >> shift) & 255u;
#pragma unroll
    for (int q = 0; q < kSortWaves; ++q) waveCount[q][t] = 0u;
    __syncthreads();
    // lanes of this wave holding the same digit
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (digit >> bit) & 1u;
      const unsigned long long m = __ballot(valid && set);
      peers &= set ? m : ~m;
    }
    const unsigned rank = static_cast<unsigned>(__popcll(peers & below));
    if (valid && rank == 0u) waveCount[w][digit] = static_cast<unsigned>(__popcll(peers));
    __syncthreads();
    {  // per digit: offsets of this round's waves after the earlier rounds
      unsigned run = running[t];
#pragma unroll
      for (int q = 0; q < kSortWaves; ++q) {
        const unsigned c = waveCount[q][t];
        waveCount[q][t] = run;
        run += c;
      }
      running[t] = run;
    }
    __syncthreads();
    if (valid) {
      const unsigned dest = start[digit] + waveCount[w][digit] + rank;
      keyOut[dest] = k;
      valOut[dest] = v;
    }
    __syncthreads();
  }
}

// Items per workgroup: 256 x 2^r, grown with n so that the digit table
// stays near 256 x 2048 entries (at most 16 rounds per workgroup).
int sortItems(unsigned n) {
  int rounds = 1;
  while (rounds < 16 && static_cast<unsigned long long>(n) > 256ull * rounds * 2048ull) rounds *= 2;
  return kSortThreads * rounds;
}

struct SortLayout {
  int items, blocks;
  std::size_t table, scanBytes, bytes;
};

SortLayout sortLayout(unsigned n) {
  SortLayout L{};
  L.items = sortItems(n);
  L.blocks = static_cast<int>((static_cast<unsigned long long>(n) + L.items - 1) / L.items);
  L.table = static_cast<std::size_t>(256) * std::max(L.blocks, 1);
  L.scanBytes = scanTempBytes32(L.table);
  // histogram table, its scan, scan scratch, one key and one value buffer
  L.bytes = (2 * L.table + 2 * static_cast<std::size_t>(n)) * sizeof(unsigned) + L.scanBytes + 256;
  return L;
}

}  // namespace

std::size_t scanTempBytes32(std::size_t n) {
  const std::size_t tiles = scanTiles<unsigned>(n);
  return tiles > 1 ? tiles * sizeof(unsigned) : 0;
}
std::size_t scanTempBytes64(std::size_t n) {
  const std::size_t tiles = scanTiles<unsigned long long>(n);
  return tiles > 1 ? tiles * sizeof(unsigned long long) : 0;
}
void exclusiveScan32(const unsigned* in, unsigned* out, std::size_t n, void* temp, std::size_t bytes,
                     hipStream_t st) {
  exclusiveScanImpl<unsigned>(in, out, n, temp, bytes, st);
}
void exclusiveScan64(const unsigned long long* in, unsigned long long* out, std::size_t n, void* temp,
                     std::size_t bytes, hipStream_t st) {
  exclusiveScanImpl<unsigned long long>(in, out, n, temp, bytes, st);
}

void tileSort(void* temp, std::size_t& bytes, unsigned* key, unsigned* keySorted, unsigned* perm,
              unsigned* permSorted, unsigned n, int bits, hipStream_t st) {
  if (n >= (1u << 31)) throw std::invalid_argument("Tile sort: more than 2^31 - 1 entries");
  const SortLayout L = sortLayout(n);
  if (!temp) {
    bytes = L.bytes;
    return;
  }
  if (bytes < L.bytes) throw std::logic_error("Tile sort scratch too small");
  if (n == 0) return;
  unsigned* hist = static_cast<unsigned*>(temp);
  unsigned* start = hist + L.table;
  unsigned* keyMid = start + L.table;
  unsigned* valMid = keyMid + n;
  void* scanTemp = valMid + n;
  const int passes = std::max(1, (std::min(bits, 32) + 7) / 8);
  // The last pass writes the sorted buffers; earlier passes alternate so
  // that no pass reads and writes the same buffer.
  const unsigned* kin = key;
  const unsigned* vin = perm;
  for (int pass = 0; pass < passes; ++pass) {
    const bool toSorted = ((passes - 1 - pass) & 1) == 0;
    unsigned* kout = toSorted ? keySorted : keyMid;
    unsigned* vout = toSorted ? permSorted : valMid;
    const int shift = 8 * pass;
    kDigitHistogram<<<L.blocks, kSortThreads, 0, st>>>(kin, static_cast<int>(n), shift, L.items, hist, L.blocks);
    launchCheck("tile sort histogram");
    exclusiveScan32(hist, start, L.table, scanTemp, L.scanBytes, st);
    kDigitScatter<<<L.blocks, kSortThreads, 0, st>>>(kin, vin, kout, vout, static_cast<int>(n), shift, L.items,
                                                     start, L.blocks);
    launchCheck("tile sort scatter");
    kin = kout;
    vin = vout;
  }
}

}  // namespace aprl_gpu

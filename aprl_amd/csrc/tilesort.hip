// Stable counting sort of (tile key, particle index) pairs for keys of at
// most 16 bits: one or two least-significant-digit passes over 8-bit
// digits. Each pass is a per-block digit histogram, one exclusive scan of
// the (digit, block) table, and a scatter whose in-block ranks come from
// wavefront ballots (lanes of one digit found with eight ballots, their
// rank by popcount of the lower lanes) and a per-digit prefix over the
// block's waves. Ranks follow the input order, so the sort is stable and
// its permutation independent of timing. Wider keys use rocprim's radix
// sort.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "particles.hpp"

namespace aprl_gpu {

constexpr int kSortThreads = 256;  // one item per thread and round
constexpr int kSortWaves = kSortThreads / 64;

// Digit histogram of `items` consecutive keys per block; hist is
// digit-major: hist[digit * blocks + block].
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
  hist[threadIdx.x * blocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one digit pass. Block b owns input items
// [b * items, (b + 1) * items), processed in rounds of kSortThreads in
// input order. The output position of the block's first item of digit d is
// the count of all items with a smaller digit plus the count of digit d in
// earlier blocks, both read from the (digit, block) histogram table.
__global__ void __launch_bounds__(kSortThreads) kDigitScatter(const unsigned* keyIn, const unsigned* valIn,
                                                              unsigned* keyOut, unsigned* valOut, int n, int shift,
                                                              int items, const unsigned* hist, int blocks) {
  __shared__ unsigned waveCount[kSortWaves][256];
  __shared__ unsigned running[256];
  __shared__ unsigned start[256];  // output position of the block's first item per digit
  const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
  {
    const unsigned* row = hist + static_cast<size_t>(t) * blocks;
    unsigned total = 0u, before = 0u;
#pragma unroll 8
    for (int q = 0; q < blocks; ++q) {
      const unsigned c = row[q];
      total += c;
      before += q < static_cast<int>(blockIdx.x) ? c : 0u;
    }
    // exclusive scan of the digit totals (Hillis-Steele over 256 lanes)
    start[t] = total;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const unsigned add = t >= off ? start[t - off] : 0u;
      __syncthreads();
      start[t] += add;
      __syncthreads();
    }
    const unsigned digitStart = start[t] - total;
    __syncthreads();
    start[t] = digitStart + before;
    running[t] = 0u;
  }
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
    {  // per digit: offsets of the waves of this round after the earlier rounds
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

// Items per block: at most 128 blocks, so that every scatter block can read
// its column of the (digit, block) table. Above
// kCountingSortLimit items the rounds per block grow too long and rocprim's
// onesweep radix sort takes over.
constexpr unsigned kCountingSortLimit = 1u << 18;
static int itemsPerBlock(unsigned n) {
  int items = kSortThreads;
  while ((static_cast<unsigned long long>(n) + items - 1) / items > 128ull) items *= 2;
  return items;
}

using WideSortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                  rocprim::default_config, 0>;

// Scratch layout for the counting sort: digit table, then one key and one
// value buffer for the intermediate pass.
void tileSort(void* temp, std::size_t& bytes, unsigned* key, unsigned* keySorted, unsigned* perm,
              unsigned* permSorted, unsigned n, int bits, hipStream_t st) {
  if (bits > 16 || n > kCountingSortLimit) {
    hipCheck(rocprim::radix_sort_pairs<WideSortConfig>(temp, bytes, key, keySorted, perm, permSorted, n, 0,
                                                       static_cast<unsigned>(bits), st),
             temp ? "radix_sort_pairs" : "radix_sort_pairs size");
    return;
  }
  const int items = itemsPerBlock(n);
  const int blocks = static_cast<int>((static_cast<unsigned long long>(n) + items - 1) / items);
  const std::size_t table = static_cast<std::size_t>(256) * std::max(blocks, 1);
  const std::size_t need = (table + 2 * static_cast<std::size_t>(n)) * sizeof(unsigned) + 256;
  if (!temp) {
    // also size a later rocprim fallback for wide keys
    std::size_t wide = 0;
    hipCheck(rocprim::radix_sort_pairs<WideSortConfig>(nullptr, wide, key, keySorted, perm, permSorted, n, 0, 32u),
             "radix_sort_pairs size");
    bytes = std::max(need, wide);
    return;
  }
  if (bytes < need) throw std::logic_error("Tile sort scratch too small");
  if (n == 0) return;
  unsigned* hist = static_cast<unsigned*>(temp);
  unsigned* keyMid = hist + table;
  unsigned* valMid = keyMid + n;
  const int passes = bits > 8 ? 2 : 1;
  const unsigned* kin = key;
  const unsigned* vin = perm;
  for (int pass = 0; pass < passes; ++pass) {
    const bool last = pass + 1 == passes;
    unsigned* kout = last ? keySorted : keyMid;
    unsigned* vout = last ? permSorted : valMid;
    const int shift = 8 * pass;
    kDigitHistogram<<<blocks, kSortThreads, 0, st>>>(kin, static_cast<int>(n), shift, items, hist, blocks);
    kDigitScatter<<<blocks, kSortThreads, 0, st>>>(kin, vin, kout, vout, static_cast<int>(n), shift, items, hist,
                                                   blocks);
    launchCheck("tile sort");
    kin = kout;
    vin = vout;
  }
}

}  // namespace aprl_gpu

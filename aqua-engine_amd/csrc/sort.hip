// Spatial ordering of radiance-cache queries.
//
// Queries leave the shading kernel in path order, which at the query vertex
// is spatially random, so each wave's hash-grid gathers touch 16 x 8 x 16
// unrelated cache lines. Sorting the queries by the Morton code of their
// normalised position before inference makes a wave's rows (and the rows of
// the waves an XCD runs together) neighbours in space, so the coarse and
// middle levels of the grid hit in L2. (Sorting on fewer key bits costs more
// in inference than it saves in the sort: with the top 12 bits only, a 4-spp
// 1080p frame's cache work took 3.1 ms instead of 2.2.)
//
// Keys: 10 bits per axis interleaved (30 bits); bits [kSortLowBit, 30) are
// sorted, cells of 1/256 of the scene per axis. Only the real queries
// [0, *count) are sorted (the count stays on the device); the permutation's
// entries past it are not written.
//
// The sort is a stable least-significant-digit radix sort, three passes of 8
// bits, each of three kernels over tiles of 4096 keys: per-tile digit counts
// (digit-major), a scan of each digit's counts across the tiles, and a stable
// scatter. In the scatter each wave ranks its own 512 keys of the tile by
// digit (eight ballots select the lanes that share a lane's digit; a running
// count per digit in LDS orders the wave's rounds), the waves are placed by a
// scan of their counts per digit, the tile is laid out in LDS in digit order,
// and each digit's run of keys is written out contiguously (a run per digit
// and tile instead of a scattered word per key). Nothing waits on another
// block, so no launch order or residency is assumed. Stable like any LSD
// sort, it orders equal keys by slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device/morton.h"

namespace aqua {

constexpr unsigned kSortLowBit = 6;
constexpr unsigned kSortPasses = 3;  // 8-bit digits over bits [6, 30)
constexpr uint32_t kBins = 256;
constexpr uint32_t kSortThreads = 512;
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kTileRounds = 8;
constexpr uint32_t kTile = kSortThreads * kTileRounds;  // keys per tile
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanChunk = kScanThreads * 16;  // counts per chunk of the scan

namespace {

uint32_t sort_tiles(uint32_t n) { return uint32_t((uint64_t(n) + kTile - 1) / kTile); }

__device__ __forceinline__ uint32_t real_count(const uint32_t* count, uint32_t n) {
  return count ? min(*count, n) : n;
}

// rec: 4 float4 per row, row 0's xyz = position in [0,1]^3.
__global__ void k_query_keys(const float4* rec, uint32_t n, uint32_t* key) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = rec[4 * size_t(i)];
    key[i] = aqd::morton30(p.x, p.y, p.z);
  }
}

// Exclusive prefix sum over the first 256 threads of the block (one value
// each; every thread of the block calls it, threads past 256 get 0).
__device__ uint32_t block_excl_scan256(uint32_t v, uint32_t* lds) {
  const uint32_t t = threadIdx.x;
  if (t < 256u) lds[t] = v;
  __syncthreads();
#pragma unroll
  for (uint32_t off = 1; off < 256u; off <<= 1) {
    const uint32_t add = (t < 256u && t >= off) ? lds[t - off] : 0u;
    __syncthreads();
    if (t < 256u) lds[t] += add;
    __syncthreads();
  }
  return t < 256u ? lds[t] - v : 0u;
}

// Digit counts of tile blockIdx.x, digit-major: counts[d * tiles + tile].
// Tiles past the real count write zeros. Each wave counts into its own
// histogram.
__global__ void __launch_bounds__(kSortThreads) k_sort_count(const uint32_t* keys, const uint32_t* count,
                                                            uint32_t n, uint32_t shift, uint32_t tiles,
                                                            uint32_t* counts) {
  __shared__ uint32_t hist[kSortWaves][kBins];
  const uint32_t m = real_count(count, n);
  for (uint32_t e = threadIdx.x; e < kSortWaves * kBins; e += kSortThreads) (&hist[0][0])[e] = 0u;
  __syncthreads();
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t t0 = blockIdx.x * kTile;
  uint32_t k[kTileRounds];  // every load in flight before the first atomic
#pragma unroll
  for (uint32_t j = 0; j < kTileRounds; ++j) {
    const uint32_t i = t0 + j * kSortThreads + threadIdx.x;
    k[j] = i < m ? keys[i] : 0xFFFFFFFFu;
  }
#pragma unroll
  for (uint32_t j = 0; j < kTileRounds; ++j)
    if (t0 + j * kSortThreads + threadIdx.x < m) atomicAdd(&hist[w][(k[j] >> shift) & (kBins - 1u)], 1u);
  __syncthreads();
  if (threadIdx.x < kBins) {
    uint32_t c = 0u;
#pragma unroll
    for (uint32_t q = 0; q < kSortWaves; ++q) c += hist[q][threadIdx.x];
    counts[size_t(threadIdx.x) * tiles + blockIdx.x] = c;
  }
}

// Block d: digit d's counts across the tiles become their exclusive prefix
// (in place); totals[d] = the digit's count over all tiles. The row goes
// through LDS in chunks, loaded and stored coalesced.
__global__ void __launch_bounds__(kScanThreads) k_sort_scan(uint32_t* counts, uint32_t tiles, uint32_t* totals) {
  __shared__ uint32_t row[kScanChunk];
  __shared__ uint32_t lds[kScanThreads];
  __shared__ uint32_t chunk_end;
  uint32_t* c = counts + size_t(blockIdx.x) * tiles;
  const uint32_t t = threadIdx.x;
  constexpr uint32_t per = kScanChunk / kScanThreads;
  uint32_t carry = 0u;
  for (uint32_t base = 0; base < tiles; base += kScanChunk) {
    const uint32_t len = min(kScanChunk, tiles - base);
    for (uint32_t e = t; e < kScanChunk; e += kScanThreads) row[e] = e < len ? c[base + e] : 0u;
    __syncthreads();
    uint32_t sum = 0u;
#pragma unroll
    for (uint32_t k = 0; k < per; ++k) sum += row[t * per + k];
    uint32_t run = carry + block_excl_scan256(sum, lds);
#pragma unroll
    for (uint32_t k = 0; k < per; ++k) {
      const uint32_t v = row[t * per + k];
      row[t * per + k] = run;
      run += v;
    }
    if (t == kScanThreads - 1u) chunk_end = run;  // the last thread's running sum
    __syncthreads();
    carry = chunk_end;
    for (uint32_t e = t; e < len; e += kScanThreads) c[base + e] = row[e];
    __syncthreads();
  }
  if (t == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of tile blockIdx.x by the digit at `shift`: key i goes to
// (keys of smaller digits) + (keys of its digit in earlier tiles) + (keys of
// its digit before it in the tile). Values are vin[i], or i when vin is null;
// keys are written only when kout is not null (the last pass needs only the
// values). Wave w owns keys [64 kTileRounds w, 64 kTileRounds (w + 1)) of the
// tile and ranks them round by round against its own running digit counts
// (one wave's LDS accesses complete in program order, so no barrier is
// needed); one pass over the waves' counts per digit then places the waves.
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(const uint32_t* kin, const uint32_t* vin,
                                                              const uint32_t* count, uint32_t n, uint32_t shift,
                                                              uint32_t tiles, const uint32_t* counts,
                                                              const uint32_t* totals, uint32_t* kout,
                                                              uint32_t* vout) {
  __shared__ uint32_t skey[kTile], sval[kTile];  // the tile in digit order
  __shared__ uint32_t scan[kBins];
  __shared__ uint32_t gbase[kBins];   // global position of the tile's first key of digit d
  __shared__ uint32_t lstart[kBins];  // its position in the tile
  __shared__ uint32_t wrun[kSortWaves][kBins];  // keys of digit d in wave w's segment (then: where they start)
  const uint32_t m = real_count(count, n);
  const uint32_t t0 = blockIdx.x * kTile;
  if (t0 >= m) return;  // the whole block
  const uint32_t len = min(kTile, m - t0);
  const uint32_t d = threadIdx.x;  // digit, for the first 256 threads
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t seg = w * (64u * kTileRounds);  // the wave's first key in the tile
  uint32_t keys[kTileRounds], vals[kTileRounds];  // every load in flight before the ranking
#pragma unroll
  for (uint32_t j = 0; j < kTileRounds; ++j) {
    const uint32_t r = seg + 64u * j + lane;
    keys[j] = r < len ? kin[t0 + r] : 0u;
    vals[j] = r < len ? (vin ? vin[t0 + r] : t0 + r) : 0u;
  }
  uint32_t in_tile = 0u, before = 0u, total = 0u;
  if (d < kBins) {
    const uint32_t* cd = counts + size_t(d) * tiles;
    before = cd[blockIdx.x];
    total = totals[d];
    in_tile = (blockIdx.x + 1u < tiles ? cd[blockIdx.x + 1u] : total) - before;
  }
  for (uint32_t e = threadIdx.x; e < kSortWaves * kBins; e += kSortThreads) (&wrun[0][0])[e] = 0u;
  const uint32_t digit_base = block_excl_scan256(total, scan);  // (its barriers also cover wrun's zeroing)
  const uint32_t local = block_excl_scan256(in_tile, scan);
  if (d < kBins) {
    gbase[d] = digit_base + before;
    lstart[d] = local;
  }
  // rank within the wave's segment: lanes sharing a digit in a round, after
  // the digit's keys of the segment's earlier rounds
  uint32_t rank[kTileRounds];
#pragma unroll
  for (uint32_t j = 0; j < kTileRounds; ++j) {
    const bool ok = seg + 64u * j + lane < len;
    const uint32_t dg = (keys[j] >> shift) & (kBins - 1u);
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const bool bit = (dg >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below =
        __builtin_amdgcn_mbcnt_hi(uint32_t(peers >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(peers), 0u));
    const uint32_t run = ok ? wrun[w][dg] : 0u;
    rank[j] = run + below;
    // the lowest lane of each digit's lanes advances the count (after every
    // lane of the wave has read it: the read above is issued first)
    __builtin_amdgcn_wave_barrier();
    if (ok && below == 0u) wrun[w][dg] = run + uint32_t(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // digit thread: where each wave's keys of the digit start in the tile
  if (d < kBins) {
    uint32_t next = local;
#pragma unroll
    for (uint32_t q = 0; q < kSortWaves; ++q) {
      const uint32_t c = wrun[q][d];
      wrun[q][d] = next;
      next += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kTileRounds; ++j) {
    if (seg + 64u * j + lane < len) {
      const uint32_t pos = wrun[w][(keys[j] >> shift) & (kBins - 1u)] + rank[j];
      skey[pos] = keys[j];
      sval[pos] = vals[j];
    }
  }
  __syncthreads();
  // each digit's run of the tile to its place, consecutive threads to
  // consecutive addresses
  for (uint32_t p = threadIdx.x; p < len; p += kSortThreads) {
    const uint32_t key = skey[p];
    const uint32_t dg = (key >> shift) & (kBins - 1u);
    const uint32_t g = gbase[dg] + (p - lstart[dg]);
    if (kout) kout[g] = key;
    vout[g] = sval[p];
  }
}

}  // namespace

// Scratch the sort needs for n rows: the per-tile digit counts and the
// digit totals.
hipError_t query_sort_storage(uint32_t n, size_t* bytes) {
  *bytes = sizeof(uint32_t) * (size_t(kBins) * sort_tiles(n) + kBins);
  return hipSuccess;
}

// Sorts the first n query slots by Morton key, or the first *d_count of them
// (a device count; null means all n are real). keys and idx hold 2 n words
// each (double buffers); *perm receives the sorted slot order, its first
// *d_count entries (the rest are not written). temp holds temp_bytes of at
// least query_sort_storage(n). With rec == nullptr the producer has already
// written keys[i] = Morton key for the real queries [0, *d_count) (d_count
// required); otherwise the keys are made from rec here. Stream-ordered, no
// host sync.
hipError_t query_sort(void* temp, size_t temp_bytes, const float4* rec, const uint32_t* d_count, uint32_t n,
                      uint32_t* keys, uint32_t* idx, hipStream_t s, const uint32_t** perm) {
  size_t need = 0;
  if (hipError_t e = query_sort_storage(n, &need); e != hipSuccess) return e;
  if (n == 0 || temp_bytes < need || (!rec && !d_count)) return hipErrorInvalidValue;
  if (rec) {
    const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 8192u);
    k_query_keys<<<blocks, 256, 0, s>>>(rec, n, keys);
  }
  const uint32_t tiles = sort_tiles(n);
  uint32_t* counts = static_cast<uint32_t*>(temp);
  uint32_t* totals = counts + size_t(kBins) * tiles;
  // pass p reads half p % 2 of keys / idx and writes the other half; the
  // first pass takes the values as the slots themselves, the last writes no
  // keys
  for (unsigned p = 0; p < kSortPasses; ++p) {
    const uint32_t shift = kSortLowBit + 8u * p;
    const uint32_t* kin = keys + (p & 1u) * size_t(n);
    const uint32_t* vin = p == 0 ? nullptr : idx + (p & 1u) * size_t(n);
    uint32_t* kout = p + 1 == kSortPasses ? nullptr : keys + ((p + 1u) & 1u) * size_t(n);
    uint32_t* vout = idx + ((p + 1u) & 1u) * size_t(n);
    k_sort_count<<<tiles, kSortThreads, 0, s>>>(kin, d_count, n, shift, tiles, counts);
    k_sort_scan<<<kBins, kScanThreads, 0, s>>>(counts, tiles, totals);
    k_sort_scatter<<<tiles, kSortThreads, 0, s>>>(kin, vin, d_count, n, shift, tiles, counts, totals, kout, vout);
  }
  *perm = idx + (kSortPasses & 1u) * size_t(n);
  return hipGetLastError();
}

}  // namespace aqua

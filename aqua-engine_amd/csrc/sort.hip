// Spatial ordering of radiance-cache queries.
//
// Queries leave the shading kernel in path order, which at the query vertex
// is spatially random, so each wave's hash-grid gathers touch 16 x 8 x 16
// unrelated cache lines. Sorting the queries by the Morton code of their
// normalised position before inference makes a wave's rows (and the rows of
// the waves an XCD runs together) neighbours in space, so the coarse and
// middle levels of the grid hit in L2.
//
// Keys: 10 bits per axis interleaved (30 bits); rows at or past the device-
// side query count get the all-ones key and sort last. Only key bits
// [kSortLowBit, 30) are sorted: three 8-bit radix passes, cells of 1/256 of
// the scene per axis.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cstdint>

#include "device/morton.h"

namespace aqua {

constexpr unsigned kSortLowBit = 6;
constexpr unsigned kSortHighBit = 30;

namespace {

// rec: 4 float4 per row, row 0's xyz = position in [0,1]^3.
__global__ void k_query_keys(const float4* rec, const uint32_t* count, uint32_t n, uint32_t* key,
                             uint32_t* idx) {
  const uint32_t m = count ? *count : n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = 0xFFFFFFFFu;
    if (i < m) {
      float4 p = rec[4 * size_t(i)];
      k = aqd::morton30(p.x, p.y, p.z);
    }
    key[i] = k;
    idx[i] = i;
  }
}

// Slots [*count, n) of keys the producer already wrote for [0, *count):
// the all-ones key (sorts last) and their own index.
__global__ void k_query_tail(const uint32_t* count, uint32_t n, uint32_t* key, uint32_t* idx) {
  const uint32_t m = *count;
  for (uint32_t i = m + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    key[i] = 0xFFFFFFFFu;
    idx[i] = i;
  }
}

}  // namespace

// Temporary storage rocPRIM needs to sort n pairs.
hipError_t query_sort_storage(uint32_t n, size_t* bytes) {
  rocprim::double_buffer<uint32_t> kb(nullptr, nullptr), vb(nullptr, nullptr);
  *bytes = 0;
  return rocprim::radix_sort_pairs(nullptr, *bytes, kb, vb, n, kSortLowBit, kSortHighBit);
}

// Sorts the first n query slots by Morton key (slots at or past *d_count
// last; a null d_count means all n are real). keys and idx hold 2 n words
// each (double buffers); *perm receives the sorted slot order (n entries,
// the real queries first: LSD radix sort is stable, so they precede the
// padding even where a key saturates). temp holds temp_bytes of at least
// query_sort_storage(n). With rec == nullptr the producer has already
// written keys[i] = Morton key and idx[i] = i for the real queries
// [0, *d_count) (d_count required); only the padding slots are filled here.
// Stream-ordered, no host sync.
hipError_t query_sort(void* temp, size_t temp_bytes, const float4* rec, const uint32_t* d_count, uint32_t n,
                      uint32_t* keys, uint32_t* idx, hipStream_t s, const uint32_t** perm) {
  size_t blocks = (size_t(n) + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  if (rec)
    k_query_keys<<<int(blocks), 256, 0, s>>>(rec, d_count, n, keys, idx);
  else if (d_count)
    k_query_tail<<<int(blocks), 256, 0, s>>>(d_count, n, keys, idx);
  else
    return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  rocprim::double_buffer<uint32_t> kb(keys, keys + n), vb(idx, idx + n);
  e = rocprim::radix_sort_pairs(temp, temp_bytes, kb, vb, n, kSortLowBit, kSortHighBit, s);
  *perm = vb.current();
  return e;
}

}  // namespace aqua

// Device-wide exclusive scan and stable LSD radix sort used by the CSR
// compaction and the first-occurrence deduplication of the hot path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rin {

class DevMem;

// out[i] = sum(in[0..i)), returns nothing; *total (device, optional) = sum.
// `in` and `out` may alias.  Works in place on up to 2^32-1 total.
void scan_u32(DevMem& mem, const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, hipStream_t s);

// Stable sort of (key, val) pairs by key bits [0, bits).  Results are left in
// keys/vals.  Uses scratch from `mem`.
void sort_pairs(DevMem& mem, uint64_t* keys, uint32_t* vals, size_t n, int bits, hipStream_t s);

// Sort an index permutation by a 128-bit key given as (hi, lo) arrays:
// perm[i] ends up listing indices in ascending (hi, lo) order, stable.
void sort_index_by2(DevMem& mem, const uint64_t* hi, const uint64_t* lo, int hi_bits, int lo_bits, uint32_t* perm,
                    size_t n, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);
void iota_u32(uint32_t* p, size_t n, hipStream_t s);

}  // namespace rin

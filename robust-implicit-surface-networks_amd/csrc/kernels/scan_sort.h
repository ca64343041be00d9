// Device-wide exclusive scan and stable LSD radix sort used by the CSR
// compaction and the first-occurrence deduplication of the hot path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rin {

class DevMem;

// An element count known either on the host or, for captured launches, only
// in device memory (then `cap` bounds it and sizes the launch).
struct Count {
    uint64_t v = 0;
    const uint32_t* d = nullptr;
    uint64_t cap = 0;
    const uint32_t* abort = nullptr;  // non-zero: an earlier stage overflowed, process nothing
    Count() = default;
    Count(uint64_t n) : v(n), cap(n) {}
    Count(const uint32_t* dev, uint64_t capacity, const uint32_t* abort_flag = nullptr)
        : d(dev), cap(capacity), abort(abort_flag) {}
    __device__ inline uint64_t get() const {
        if (!d) return v;
        if (abort && *abort) return 0;
        const uint64_t n = *d;
        return n < cap ? n : cap;
    }
};

// out[i] = sum(in[0..i)), returns nothing; *total (device, optional) = sum.
// `in` and `out` may alias.  Works in place on up to 2^32-1 total.
void scan_u32(DevMem& mem, const uint32_t* in, uint32_t* out, Count n, uint32_t* total, hipStream_t s);

// Three exclusive scans at once: out[f][i] = sum(in[f][j * stride], j < i),
// *total[f] (optional) = the sum.  Inputs and outputs must not alias.
void scan3_u32(DevMem& mem, const uint32_t* const in[3], int stride, uint32_t* const out[3], Count n,
               uint32_t* const total[3], hipStream_t s);

// Stable sort of (key, val) pairs by key bits [0, bits).  Results are left in
// keys/vals.  Uses scratch from `mem`.
void sort_pairs(DevMem& mem, uint64_t* keys, uint32_t* vals, Count n, int bits, hipStream_t s);

// Sort an index permutation by a 128-bit key given as (hi, lo) arrays:
// perm[i] ends up listing indices in ascending (hi, lo) order, stable.
void sort_index_by2(DevMem& mem, const uint64_t* hi, const uint64_t* lo, int hi_bits, int lo_bits, uint32_t* perm,
                    Count n, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);
void iota_u32(uint32_t* p, Count n, hipStream_t s);
void copy_u32(uint32_t* dst, const uint32_t* src, Count n, hipStream_t s);
void copy_u64(uint64_t* dst, const uint64_t* src, Count n, hipStream_t s);

}  // namespace rin

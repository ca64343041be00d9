// Device-wide scan and stable LSD radix sort (8-bit digits) for 64-lane
// wavefronts: blocks of 256 threads (4 waves) own tiles of 2048 elements.
// The scatter ranks equal digits inside a wave with eight ballots (one per
// digit bit), across the block's waves through LDS counters, and across
// tiles through a scanned digit-major histogram, so equal keys keep their
// input order.
#include "../devmem.h"
#include "scan_sort.h"

namespace rin {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
// the three-field scan: smaller tiles (its inputs are a few hundred
// thousand per-tet counts: four times the workgroups keep more loads in
// flight)
constexpr int kItems3 = 2;
constexpr int kTile3 = kThreads * kItems3;
constexpr int kWaves = kThreads / 64;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// exclusive scan of one value per thread across the block; returns the
// block total through `total`
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kWaves; ++w) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ void k_scan_reduce(const uint32_t* __restrict__ in, Count cn, uint32_t* __restrict__ sums) {
    const size_t n = cn.get();
    const size_t base = (size_t)blockIdx.x * kTile;
    uint32_t acc = 0;
    for (int i = 0; i < kItems; ++i) {
        size_t idx = base + (size_t)i * kThreads + threadIdx.x;
        if (idx < n) acc += in[idx];
    }
    uint32_t tot;
    (void)block_excl_scan(acc, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// offs: the tiles' scanned sums, or (direct) their raw sums, which each
// workgroup then adds up itself for the tiles before it
__global__ void k_scan_apply(const uint32_t* in, uint32_t* out, Count cn, const uint32_t* __restrict__ offs,
                             uint32_t* total, bool direct = false) {
    const size_t n = cn.get();
    uint32_t pre = 0;
    if (offs && direct) {
        uint32_t x = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kThreads) x += offs[b];
        uint32_t tot;
        (void)block_excl_scan(x, tot);
        pre = tot;
    } else if (offs) {
        pre = offs[blockIdx.x];
    }
    const size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t acc = 0;
    for (int i = 0; i < kItems; ++i) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        acc += v[i];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(acc, tot) + pre;
    for (int i = 0; i < kItems; ++i) {
        size_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    // the block holding element n-1 (or block 0 when n == 0) reports the total
    const size_t last_blk = n ? (n - 1) / kTile : 0;
    if (total && blockIdx.x == last_blk && threadIdx.x == kThreads - 1) *total = run;
}

// Three exclusive scans in one pass over in[f][i * stride], f = 0..2.
struct Scan3 {
    const uint32_t* in[3];
    uint32_t* out[3];
    uint32_t* total[3];
    int stride;
};

__global__ void k_scan3_reduce(Scan3 S, Count cn, uint32_t* __restrict__ sums, uint32_t nb) {
    const size_t n = cn.get();
    const size_t base = (size_t)blockIdx.x * kTile3;
    uint32_t acc[3] = {0, 0, 0};
    for (int i = 0; i < kItems3; ++i) {
        const size_t idx = base + (size_t)i * kThreads + threadIdx.x;
        if (idx < n)
            for (int f = 0; f < 3; ++f) acc[f] += S.in[f][idx * S.stride];
    }
    for (int f = 0; f < 3; ++f) {
        uint32_t tot;
        (void)block_excl_scan(acc[f], tot);
        if (threadIdx.x == 0) sums[f * nb + blockIdx.x] = tot;
    }
}

// one block per field: exclusive scan of that field's nb tile sums, kTile
// at a time
__global__ void k_scan3_top(uint32_t* __restrict__ sums, uint32_t nb) {
    uint32_t* p = sums + (size_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kTile) {
        const uint32_t base = b0 + threadIdx.x * kItems;
        uint32_t v[kItems], acc = 0;
        for (int i = 0; i < kItems; ++i) {
            v[i] = base + i < nb ? p[base + i] : 0u;
            acc += v[i];
        }
        uint32_t tot;
        uint32_t run = carry + block_excl_scan(acc, tot);
        for (int i = 0; i < kItems; ++i) {
            if (base + i < nb) p[base + i] = run;
            run += v[i];
        }
        carry += tot;
    }
}

__global__ void k_scan3_apply(Scan3 S, Count cn, const uint32_t* __restrict__ sums, uint32_t nb) {
    const size_t n = cn.get();
    const size_t base = (size_t)blockIdx.x * kTile3 + (size_t)threadIdx.x * kItems3;
    const size_t last_blk = n ? (n - 1) / kTile3 : 0;
    for (int f = 0; f < 3; ++f) {
        uint32_t v[kItems3], acc = 0;
        for (int i = 0; i < kItems3; ++i) {
            const size_t idx = base + i;
            v[i] = idx < n ? S.in[f][idx * S.stride] : 0u;
            acc += v[i];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan(acc, tot) + sums[f * nb + blockIdx.x];
        for (int i = 0; i < kItems3; ++i) {
            const size_t idx = base + i;
            if (idx < n) S.out[f][idx] = run;
            run += v[i];
        }
        if (S.total[f] && blockIdx.x == last_blk && threadIdx.x == kThreads - 1) *S.total[f] = run;
    }
}

// The same without the top-level scan: each workgroup adds up the tile sums
// before its own (a few hundred tiles: at most a few loads per lane)
__global__ void k_scan3_apply_direct(Scan3 S, Count cn, const uint32_t* __restrict__ sums, uint32_t nb) {
    const size_t n = cn.get();
    uint32_t pre[3] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kThreads)
        for (int f = 0; f < 3; ++f) pre[f] += sums[f * nb + b];
    for (int f = 0; f < 3; ++f) {
        uint32_t tot;
        (void)block_excl_scan(pre[f], tot);
        pre[f] = tot;
    }
    const size_t base = (size_t)blockIdx.x * kTile3 + (size_t)threadIdx.x * kItems3;
    const size_t last_blk = n ? (n - 1) / kTile3 : 0;
    for (int f = 0; f < 3; ++f) {
        uint32_t v[kItems3], acc = 0;
        for (int i = 0; i < kItems3; ++i) {
            const size_t idx = base + i;
            v[i] = idx < n ? S.in[f][idx * S.stride] : 0u;
            acc += v[i];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan(acc, tot) + pre[f];
        for (int i = 0; i < kItems3; ++i) {
            const size_t idx = base + i;
            if (idx < n) S.out[f][idx] = run;
            run += v[i];
        }
        if (S.total[f] && blockIdx.x == last_blk && threadIdx.x == kThreads - 1) *S.total[f] = run;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, Count cn) {
    const size_t n = cn.get();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}
__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t* dst,
                             Count cn) {
    const size_t n = cn.get();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

__global__ void k_copy_u32(uint32_t* __restrict__ d, const uint32_t* __restrict__ s, Count cn) {
    const size_t n = cn.get();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}
__global__ void k_copy_u64(uint64_t* __restrict__ d, const uint64_t* __restrict__ s, Count cn) {
    const size_t n = cn.get();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}

__global__ void k_rs_hist(const uint64_t* __restrict__ keys, Count cn, int shift, uint32_t* __restrict__ hist,
                          uint32_t nblocks) {
    const size_t n = cn.get();
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
    for (int r = 0; r < kItems; ++r) {
        size_t i = base + (size_t)r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                             uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, Count cn, int shift,
                             const uint32_t* __restrict__ offs, uint32_t nblocks) {
    const size_t n = cn.get();
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kWaves][256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = offs[(size_t)t * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kItems; ++r) {
        const size_t i = base + (size_t)r * kThreads + t;
        const bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t same = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            same &= bit ? bb : ~bb;
        }
        const uint32_t lower = __popcll(same & lt);
        if (valid && lower == 0) wcnt[wave][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + lower;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

inline unsigned grid_for(size_t n, unsigned threads = 256) {
    size_t g = (n + threads - 1) / threads;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)(g ? g : 1);
}

}  // namespace

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}
void fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}
void iota_u32(uint32_t* p, Count n, hipStream_t s) {
    if (n.cap) hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n.cap)), dim3(256), 0, s, p, n);
}
void copy_u32(uint32_t* d, const uint32_t* src, Count n, hipStream_t s) {
    if (n.cap) hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n.cap)), dim3(256), 0, s, d, src, n);
}
void copy_u64(uint64_t* d, const uint64_t* src, Count n, hipStream_t s) {
    if (n.cap) hipLaunchKernelGGL(k_copy_u64, dim3(grid_for(n.cap)), dim3(256), 0, s, d, src, n);
}

void scan_u32(DevMem& mem, const uint32_t* in, uint32_t* out, Count n, uint32_t* total, hipStream_t s) {
    if (n.cap == 0) {
        if (total) fill_u32(total, 0, 1, s);
        return;
    }
    const size_t nb = (n.cap + kTile - 1) / kTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kThreads), 0, s, in, out, n, (const uint32_t*)nullptr, total);
        return;
    }
    DBuf<uint32_t> sums(mem, nb);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kThreads), 0, s, in, n, sums.p);
    if (nb <= 4096) {
        // few tiles: each applying workgroup adds up the sums before it
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n,
                           (const uint32_t*)sums.p, total, true);
        return;
    }
    if (nb <= (size_t)16 * kTile)  // one workgroup scans the tile sums
        hipLaunchKernelGGL(k_scan3_top, dim3(1), dim3(kThreads), 0, s, sums.p, (uint32_t)nb);
    else
        scan_u32(mem, sums.p, sums.p, Count(nb), nullptr, s);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n, (const uint32_t*)sums.p,
                       total);
}

void scan3_u32(DevMem& mem, const uint32_t* const in[3], int stride, uint32_t* const out[3], Count n,
               uint32_t* const total[3], hipStream_t s) {
    const size_t nb = (n.cap + kTile3 - 1) / kTile3;
    if (nb > (size_t)16 * kTile) {  // beyond one top-level workgroup: three separate scans
        for (int f = 0; f < 3; ++f) {
            if (stride != 1) throw std::runtime_error("strided scan larger than one top-level tile");
            scan_u32(mem, in[f], out[f], n, total[f], s);
        }
        return;
    }
    Scan3 S;
    for (int f = 0; f < 3; ++f) {
        S.in[f] = in[f];
        S.out[f] = out[f];
        S.total[f] = total[f];
    }
    S.stride = stride;
    if (n.cap == 0) {
        for (int f = 0; f < 3; ++f)
            if (total[f]) fill_u32(total[f], 0, 1, s);
        return;
    }
    DBuf<uint32_t> sums(mem, 3 * nb);
    hipLaunchKernelGGL(k_scan3_reduce, dim3((unsigned)nb), dim3(kThreads), 0, s, S, n, sums.p, (uint32_t)nb);
    if (nb <= 4096) {
        // few tiles: each applying workgroup adds up the sums before it
        hipLaunchKernelGGL(k_scan3_apply_direct, dim3((unsigned)nb), dim3(kThreads), 0, s, S, n,
                           (const uint32_t*)sums.p, (uint32_t)nb);
        return;
    }
    hipLaunchKernelGGL(k_scan3_top, dim3(3), dim3(kThreads), 0, s, sums.p, (uint32_t)nb);
    hipLaunchKernelGGL(k_scan3_apply, dim3((unsigned)nb), dim3(kThreads), 0, s, S, n, (const uint32_t*)sums.p,
                       (uint32_t)nb);
}

void sort_pairs(DevMem& mem, uint64_t* keys, uint32_t* vals, Count n, int bits, hipStream_t s) {
    if (n.cap <= 1 || bits <= 0) return;
    const uint32_t nb = (uint32_t)((n.cap + kTile - 1) / kTile);
    DBuf<uint64_t> k2(mem, n.cap);
    DBuf<uint32_t> v2(mem, n.cap);
    DBuf<uint32_t> hist(mem, (size_t)256 * nb);
    uint64_t* ka = keys;
    uint32_t* va = vals;
    uint64_t* kb = k2.p;
    uint32_t* vb = v2.p;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kThreads), 0, s, ka, n, shift, hist.p, nb);
        scan_u32(mem, hist.p, hist.p, Count((size_t)256 * nb), nullptr, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kThreads), 0, s, ka, va, kb, vb, n, shift,
                           (const uint32_t*)hist.p, nb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        copy_u64(keys, ka, n, s);
        copy_u32(vals, va, n, s);
    }
}

void sort_index_by2(DevMem& mem, const uint64_t* hi, const uint64_t* lo, int hi_bits, int lo_bits, uint32_t* perm,
                    Count n, hipStream_t s) {
    iota_u32(perm, n, s);
    if (n.cap <= 1) return;
    DBuf<uint64_t> k(mem, n.cap);
    if (lo) {
        copy_u64(k.p, lo, n, s);
        sort_pairs(mem, k.p, perm, n, lo_bits, s);
    }
    hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(n.cap)), dim3(256), 0, s, hi, (const uint32_t*)perm, k.p, n);
    sort_pairs(mem, k.p, perm, n, hi_bits, s);
}

}  // namespace rin

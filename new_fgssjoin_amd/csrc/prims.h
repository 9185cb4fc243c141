// prims.h -- hand-written gfx950 device primitives used by the join pipeline:
//   * exclusive prefix scan (reduce-then-scan, any length)
//   * stable LSD radix sort of (uint32 key, uint32 index) pairs
// Wavefronts are 64 lanes; ballots are 64-bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace fgss {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// inclusive wave scan (shfl_up, 6 steps)
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T o = __shfl_up(v, d, kWave);
        if (lane_id() >= d) v += o;
    }
    return v;
}

// inclusive wave scan of 32-bit values in six DPP adds (no LDS round trips): shifts
// by 1, 2, 4 and 8 lanes inside each row of 16 (lanes shifted in from outside the
// row add 0), then lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3
__device__ __forceinline__ uint32_t wave_inclusive_scan_dpp(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

// inclusive max-scan over the wavefront (values >= 0): shifts by 1, 2, 4 and 8 lanes
// inside each row of 16 (lanes shifted in from outside the row read 0), then lane 15
// into rows 1 and 3 and lane 31 into rows 2 and 3
__device__ __forceinline__ uint32_t wave_max_scan(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));   // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// dst[i] = v for i < n (used instead of memset so that the work can be captured
// into a HIP graph)
template <class T>
__global__ void k_fill(T *__restrict__ dst, int64_t n, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

template <class T>
inline hipError_t fill(T *dst, int64_t n, T v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, st, dst, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exclusive scan
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;   // 4096 elements per block

// block-wide exclusive scan of one value per thread; returns the block total in *total
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *total)
{
    __shared__ T warp_tot[kScanThreads / kWave];
    T inc = wave_inclusive_scan(v);
    const int w = (int)threadIdx.x / kWave;
    if (lane_id() == kWave - 1) warp_tot[w] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / kWave; k++) {
        T x = warp_tot[k];
        if (k < w) base += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T *__restrict__ in, T *__restrict__ sums, int64_t n)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    T tot;
    (void)block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// scan one tile per block; each thread owns kScanItems consecutive elements
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const T *__restrict__ in, T *__restrict__ out, int64_t n,
                                                            const T *__restrict__ block_off, T *__restrict__ total_out)
{
    __shared__ T tile[kScanTile];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int idx = k * kScanThreads + (int)threadIdx.x;
        int64_t i = base + idx;
        tile[idx] = i < n ? in[i] : (T)0;
    }
    __syncthreads();
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) { v[k] = tile[threadIdx.x * kScanItems + k]; s += v[k]; }
    T tot;
    T ex = block_exclusive_scan(s, &tot);
    T off = block_off ? block_off[blockIdx.x] : (T)0;
    T run = off + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) { tile[threadIdx.x * kScanItems + k] = run; run += v[k]; }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int idx = k * kScanThreads + (int)threadIdx.x;
        int64_t i = base + idx;
        if (i < n) out[i] = tile[idx];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = off + tot;
}

// k_scan_tile whose block offset is the sum of sums[0 .. blockIdx.x), added up by the
// block itself (used when the block sums fit one tile, so that the scan of the sums
// is not a launch of its own)
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_pre(const T *__restrict__ in, T *__restrict__ out, int64_t n,
                                                                const T *__restrict__ sums, T *__restrict__ total_out)
{
    __shared__ T tile[kScanTile];
    T part = 0;
    for (int k = (int)threadIdx.x; k < (int)blockIdx.x; k += kScanThreads) part += sums[k];
    T off;
    (void)block_exclusive_scan(part, &off);
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int idx = k * kScanThreads + (int)threadIdx.x;
        int64_t i = base + idx;
        tile[idx] = i < n ? in[i] : (T)0;
    }
    __syncthreads();
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) { v[k] = tile[threadIdx.x * kScanItems + k]; s += v[k]; }
    T tot;
    T ex = block_exclusive_scan(s, &tot);
    T run = off + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) { tile[threadIdx.x * kScanItems + k] = run; run += v[k]; }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int idx = k * kScanThreads + (int)threadIdx.x;
        int64_t i = base + idx;
        if (i < n) out[i] = tile[idx];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = off + tot;
}

// Scratch needed by exclusive_scan for n elements (in elements of T).
inline int64_t scan_scratch_elems(int64_t n)
{
    int64_t tot = 0;
    while (n > kScanTile) {
        int64_t nb = (n + kScanTile - 1) / kScanTile;
        tot += nb + 1;
        n = nb;
    }
    return tot + 1;
}

// out[i] = sum(in[0..i)); *total (device pointer, may be null) = sum(in).  In-place allowed.
template <class T>
hipError_t exclusive_scan(const T *in, T *out, int64_t n, T *total, T *scratch, hipStream_t st)
{
    if (n <= 0) return total ? fill<T>(total, 1, (T)0, st) : hipSuccess;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tile<T>, dim3(1), dim3(kScanThreads), 0, st, in, out, n, (const T *)nullptr, total);
        return hipGetLastError();
    }
    T *sums = scratch;
    T *rest = scratch + nb + 1;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, sums, n);
    if (nb <= kScanTile) {   // each tile block adds up the sums before it
        hipLaunchKernelGGL(k_scan_tile_pre<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n,
                           (const T *)sums, total);
        return hipGetLastError();
    }
    hipError_t e = exclusive_scan<T>(sums, sums, nb, (T *)nullptr, rest, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_tile<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n, (const T *)sums, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, index) pairs, 8-bit digits
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortItems = 16;                        // per lane
constexpr int kSortWaveSpan = kWave * kSortItems;     // 1024 consecutive elements per wave
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 elements per block

// hist[d * ntiles + tile] = number of keys in the tile with digit d
// n_dev, when given, holds the element count (<= n, the capacity the grid was sized for)
__device__ __forceinline__ int64_t dev_count(int64_t n, const uint32_t *n_dev)
{
    return n_dev ? min(n, (int64_t)*n_dev) : n;
}

static __global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                             const uint32_t *__restrict__ n_dev, int shift,
                                                             uint32_t *__restrict__ hist, int ntiles)
{
    n = dev_count(n, n_dev);
    __shared__ uint32_t h[kSortWaves][256];
    const int w = (int)threadIdx.x / kWave;
    for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)w * kSortWaveSpan;
#pragma unroll 4
    for (int k = 0; k < kSortItems; k++) {
        int64_t i = base + (int64_t)k * kWave + lane_id();
        if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += kSortThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) s += h[k][d];
        hist[(int64_t)d * ntiles + blockIdx.x] = s;
    }
}

// Stable scatter.  offs[d * ntiles + tile] = exclusive-scanned histogram.
// If idx_in is null the index of element i is i itself.  The tile is first ranked
// into LDS in digit order (stable: waves in order, lanes ranked by a ballot match on
// the digit), then written out position by position, so that each digit's run of
// the tile goes to consecutive addresses (coalesced) instead of one scattered store
// per element.
static_assert(kSortThreads == 256, "one thread per digit");
static __global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t *__restrict__ keys_in,
                                                                const uint32_t *__restrict__ idx_in, int64_t n,
                                                                const uint32_t *__restrict__ n_dev, int shift,
                                                                const uint32_t *__restrict__ offs, int ntiles,
                                                                uint32_t *__restrict__ keys_out,
                                                                uint32_t *__restrict__ idx_out)
{
    n = dev_count(n, n_dev);
    const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
    if (tile0 >= n) return;   // the whole block: past the device-side count
    __shared__ uint32_t h[kSortWaves][256];   // per-wave digit counts, then per-wave local starts
    __shared__ uint32_t dbase[256];           // global start of (digit, tile) minus its local start
    __shared__ uint32_t skey[kSortTile], sval[kSortTile];
    const int w = (int)threadIdx.x / kWave;
    for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = tile0 + (int64_t)w * kSortWaveSpan;
    // per-wave digit counts
#pragma unroll 4
    for (int k = 0; k < kSortItems; k++) {
        int64_t i = base + (int64_t)k * kWave + lane_id();
        if (i < n) atomicAdd(&h[w][(keys_in[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    {   // digit d (this thread): its run in the tile starts at the count of smaller digits
        const int d = (int)threadIdx.x;
        uint32_t c[kSortWaves], tot = 0;
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) { c[k] = h[k][d]; tot += c[k]; }
        uint32_t all;
        const uint32_t loc = block_exclusive_scan<uint32_t>(tot, &all);
        uint32_t run = loc;
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) { h[k][d] = run; run += c[k]; }
        dbase[d] = offs[(int64_t)d * ntiles + blockIdx.x] - loc;
    }
    __syncthreads();
    // in-order ranking within each 64-element chunk (ballot match on 8 digit bits) into LDS
    for (int k = 0; k < kSortItems; k++) {
        int64_t i = base + (int64_t)k * kWave + lane_id();
        const bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bal = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t start = valid ? h[w][d] : 0u;
        // all lanes have read their base; the highest peer advances it
        if (valid && rank == cnt - 1) h[w][d] = start + cnt;
        if (valid) {
            skey[start + rank] = key;
            sval[start + rank] = idx_in ? idx_in[i] : (uint32_t)i;
        }
    }
    __syncthreads();
    // local position j of digit d goes to dbase[d] + j: consecutive within a run
    const int cnt_tile = (int)min((int64_t)kSortTile, n - tile0);
    for (int j = (int)threadIdx.x; j < cnt_tile; j += kSortThreads) {
        const uint32_t key = skey[j];
        const uint32_t dst = dbase[(key >> shift) & 255u] + (uint32_t)j;
        keys_out[dst] = key;
        idx_out[dst] = sval[j];
    }
}

inline int radix_tiles(int64_t n) { return (int)((n + kSortTile - 1) / kSortTile); }

// Scratch for radix_sort (uint32 elements): histogram + scan scratch.
inline int64_t radix_scratch_elems(int64_t n)
{
    int64_t h = (int64_t)radix_tiles(n) * 256;
    return h + scan_scratch_elems(h);
}

// Sort n (key, index) pairs by the low `bits` bits of key, stably; indices start as 0..n-1.
// Result in (keys_b, idx_b) after an odd number of passes, else (keys_a, idx_a); the
// function returns which: 0 -> a, 1 -> b.  keys_a holds the input keys.  With n_dev the
// element count is read on the device (n is then the capacity the launch is sized for).
// Stable sort of (key, value) pairs by key bits [0, bits).  Values come from vals_a
// when vals_given, else the first pass takes the identity (value of element i = i).
// The result is in (keys_b, vals_b) when *which is 1, else in (keys_a, vals_a).
inline hipError_t radix_sort_kv(uint32_t *keys_a, uint32_t *vals_a, uint32_t *keys_b, uint32_t *vals_b, int64_t n,
                                const uint32_t *n_dev, int bits, bool vals_given, uint32_t *scratch, hipStream_t st,
                                int *which)
{
    *which = 0;
    if (n <= 0) return hipSuccess;
    const int ntiles = radix_tiles(n);
    uint32_t *hist = scratch;
    uint32_t *sscr = scratch + (int64_t)ntiles * 256;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    // even passes read a and write b, odd passes read b and write a
    for (int p = 0; p < passes; p++) {
        const bool even = (p & 1) == 0;
        const uint32_t *ki = even ? keys_a : keys_b;
        const uint32_t *ii = (p == 0 && !vals_given) ? nullptr : (even ? vals_a : vals_b);
        uint32_t *ko = even ? keys_b : keys_a;
        uint32_t *io = even ? vals_b : vals_a;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kSortThreads), 0, st, ki, n, n_dev, 8 * p, hist, ntiles);
        hipError_t e = exclusive_scan<uint32_t>(hist, hist, (int64_t)ntiles * 256, nullptr, sscr, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kSortThreads), 0, st, ki, ii, n, n_dev, 8 * p, hist,
                           ntiles, ko, io);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    *which = (passes & 1) ? 1 : 0;
    return hipSuccess;
}

// Stable sort of keys with their original indices (identity values).
inline hipError_t radix_sort_pairs(uint32_t *keys_a, uint32_t *idx_a, uint32_t *keys_b, uint32_t *idx_b, int64_t n,
                                   const uint32_t *n_dev, int bits, uint32_t *scratch, hipStream_t st, int *which)
{
    return radix_sort_kv(keys_a, idx_a, keys_b, idx_b, n, n_dev, bits, false, scratch, st, which);
}

}  // namespace fgss

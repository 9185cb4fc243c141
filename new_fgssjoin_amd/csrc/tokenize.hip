// tokenize.hip -- q-gram tokenization and global-frequency token reordering on the GPU.
//
// Same result as the host tokenizer of collection.cpp (and so as the reference's
// fixture, test/data/string/small.txt -> small.txt.3gram.bin/.loc), built from device
// scans and stable radix sorts only, so that every tie is broken exactly as there:
//   1. newline positions per 4 KiB tile (count, scan, write)     io.cpp:8-26
//   2. non-empty lines, q-grams per line (max(0, len - q + 1)), scans
//   3. per q-gram: lowercase A-Z bytes packed into a key          data.cpp:36-40
//   4. occurrence number = rank among equal keys of its line: stable sort by key,
//      segment starts at (key, line) changes                      qgram.cpp:20-49
//   5. 64-bit sdbm of the bytes (stopping at a NUL) and the decimal occurrence number
//                                                                  util.cpp:146-157
//   6. stable sort by hash -> groups = distinct hashes; df = group size, first
//      occurrence = the group's smallest flat index
//   7. token id = rank by (df, first occurrence): two stable sorts of the groups
//   8. records = lines with q-grams, stable-sorted by q-gram count; internal order is
//      that order reversed; tokens sorted within each record (stable by position):
//      two stable sorts of all q-grams, by token id then by record.
// Every sort is the LSD radix sort of prims.h (8-bit digits, stable).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "fgss_internal.h"
#include "prims.h"

using namespace fgss;

#define TKCK(x)                                                                                     \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fgss_fail(FGSS_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int kTileThreads = 256;
constexpr int kTileBytes = kTileThreads * 16;   // 16 bytes per thread

__device__ __forceinline__ uint32_t to_lower(uint32_t c) { return (c >= 'A' && c <= 'Z') ? c + 32u : c; }

// the 16 bytes of a thread (fewer at the end of the text) as a bit mask of newlines
__device__ __forceinline__ uint32_t newline_mask(const uint8_t *__restrict__ text, int64_t base, int64_t nbytes)
{
    uint32_t m = 0;
    if (base + 16 <= nbytes) {
        const uint4 v = *(const uint4 *)(text + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (((w[k >> 2] >> (8 * (k & 3))) & 0xFFu) == '\n') m |= 1u << k;
    } else {
        for (int k = 0; base + k < nbytes && k < 16; k++)
            if (text[base + k] == '\n') m |= 1u << k;
    }
    return m;
}

__global__ __launch_bounds__(kTileThreads) void k_nl_count(const uint8_t *__restrict__ text, int64_t nbytes,
                                                          uint32_t *__restrict__ cnt)
{
    const int64_t base = (int64_t)blockIdx.x * kTileBytes + (int64_t)threadIdx.x * 16;
    uint32_t c = wave_sum((uint32_t)__popc(newline_mask(text, base, nbytes)));
    __shared__ uint32_t ws[kTileThreads / kWave];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < kTileThreads / kWave; k++) s += ws[k];
        cnt[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kTileThreads) void k_nl_write(const uint8_t *__restrict__ text, int64_t nbytes,
                                                          const uint32_t *__restrict__ toff,
                                                          uint32_t *__restrict__ nlpos)
{
    const int64_t base = (int64_t)blockIdx.x * kTileBytes + (int64_t)threadIdx.x * 16;
    uint32_t m = newline_mask(text, base, nbytes);
    const uint32_t c = (uint32_t)__popc(m);
    const uint32_t inc = wave_inclusive_scan(c);
    __shared__ uint32_t ws[kTileThreads / kWave];
    const int w = (int)threadIdx.x / kWave;
    if (lane_id() == kWave - 1) ws[w] = inc;
    __syncthreads();
    uint32_t off = toff[blockIdx.x] + inc - c;
    for (int k = 0; k < w; k++) off += ws[k];
    while (m) {
        const int k = __ffs(m) - 1;
        m &= m - 1;
        nlpos[off++] = (uint32_t)(base + k);
    }
}

// all lines (n_nl + 1 of them, the last one possibly empty): start, length, non-empty flag
__global__ void k_lines(const uint32_t *__restrict__ nlpos, int64_t n_nl, int64_t nbytes, uint32_t *__restrict__ lstart,
                        uint32_t *__restrict__ llen, uint32_t *__restrict__ flag)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n_nl) return;
    const int64_t s = j ? (int64_t)nlpos[j - 1] + 1 : 0;
    const int64_t e = j < n_nl ? (int64_t)nlpos[j] : nbytes;
    lstart[j] = (uint32_t)s;
    llen[j] = (uint32_t)(e - s);
    flag[j] = e > s ? 1u : 0u;
}

// non-empty line r: start, length, q-gram count
__global__ void k_lines_compact(const uint32_t *__restrict__ lstart, const uint32_t *__restrict__ llen,
                                const uint32_t *__restrict__ flag, const uint32_t *__restrict__ off, int64_t nlines,
                                int q, uint32_t *__restrict__ ls, uint32_t *__restrict__ ll, uint32_t *__restrict__ ng)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nlines || !flag[j]) return;
    const uint32_t r = off[j], L = llen[j];
    ls[r] = lstart[j];
    ll[r] = L;
    ng[r] = L >= (uint32_t)q ? L - (uint32_t)q + 1u : 0u;
}

// wave per line: the lowercase key of every q-gram (the first q - 4 bytes in hi, the
// last min(q, 4) in lo), its line and its position
__global__ __launch_bounds__(256) void k_grams(const uint8_t *__restrict__ text, const uint32_t *__restrict__ ls,
                                               const uint32_t *__restrict__ ng, const uint32_t *__restrict__ roff,
                                               int64_t nl, int q, uint32_t *__restrict__ klo,
                                               uint32_t *__restrict__ khi, uint32_t *__restrict__ gline,
                                               uint32_t *__restrict__ gpos)
{
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; r < nl; r += nwaves) {
        const uint32_t s = ls[r], n = ng[r], f0 = roff[r];
        for (uint32_t i = lane_id(); i < n; i += kWave) {
            uint32_t lo = 0, hi = 0;
            for (int j = 0; j < q; j++) {
                const uint32_t c = to_lower(text[(int64_t)s + i + j]);
                if (j < q - 4) hi = (hi << 8) | c;
                else lo = (lo << 8) | c;
            }
            klo[f0 + i] = lo;
            if (q > 4) khi[f0 + i] = hi;
            gline[f0 + i] = (uint32_t)r;
            gpos[f0 + i] = i;
        }
    }
}

__global__ void k_gather(uint32_t *__restrict__ out, const uint32_t *__restrict__ in, const uint32_t *__restrict__ perm,
                         int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

// sorted by key (perm P): a segment starts where the key or the line changes
__global__ void k_occ_flags(const uint32_t *__restrict__ P, const uint32_t *__restrict__ klo,
                            const uint32_t *__restrict__ khi, const uint32_t *__restrict__ gline, int64_t n,
                            uint32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t f = 1;
    if (i > 0) {
        const uint32_t a = P[i], b = P[i - 1];
        f = (klo[a] != klo[b] || (khi && khi[a] != khi[b]) || gline[a] != gline[b]) ? 1u : 0u;
    }
    flag[i] = f;
}

// segment id of sorted position i from the exclusive scan ex of the start flags
__device__ __forceinline__ uint32_t seg_of(const uint32_t *flag, const uint32_t *ex, int64_t i)
{
    return ex[i] + flag[i] - 1u;
}

__global__ void k_seg_start(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex, int64_t n,
                            uint32_t *__restrict__ start)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) start[ex[i]] = (uint32_t)i;
}

__global__ void k_occ(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex,
                      const uint32_t *__restrict__ start, const uint32_t *__restrict__ P, int64_t n,
                      uint32_t *__restrict__ occ)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) occ[P[i]] = (uint32_t)i - start[seg_of(flag, ex, i)];
}

// sdbm over the lowercase q-gram bytes, stopping at a NUL, then over the decimal
// digits of the occurrence number
__global__ void k_hash(const uint8_t *__restrict__ text, const uint32_t *__restrict__ ls,
                       const uint32_t *__restrict__ gline, const uint32_t *__restrict__ gpos,
                       const uint32_t *__restrict__ occ, int64_t n, int q, uint32_t *__restrict__ hlo,
                       uint32_t *__restrict__ hhi)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int64_t s = (int64_t)ls[gline[f]] + gpos[f];
    uint64_t h = 0;
    bool nul = false;
    for (int j = 0; j < q; j++) {
        const uint32_t c = to_lower(text[s + j]);
        if (c == 0) { nul = true; break; }
        h = (uint64_t)c + (h << 6) + (h << 16) - h;
    }
    if (!nul) {
        const uint32_t o = occ[f];
        uint32_t p10 = 1;
        while (o / p10 >= 10u) p10 *= 10u;
        for (; p10 > 0; p10 /= 10u) h = (uint64_t)('0' + (o / p10) % 10u) + (h << 6) + (h << 16) - h;
    }
    hlo[f] = (uint32_t)h;
    hhi[f] = (uint32_t)(h >> 32);
}

// sorted by hash (perm P, sorted high words shi): group starts where the hash changes
__global__ void k_hash_flags(const uint32_t *__restrict__ P, const uint32_t *__restrict__ hlo,
                             const uint32_t *__restrict__ shi, int64_t n, uint32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || shi[i] != shi[i - 1] || hlo[P[i]] != hlo[P[i - 1]]) ? 1u : 0u;
}

__global__ void k_group_info(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex,
                             const uint32_t *__restrict__ P, const uint32_t *__restrict__ hlo,
                             const uint32_t *__restrict__ shi, int64_t n, uint32_t *__restrict__ gstart,
                             uint32_t *__restrict__ gfirst, uint32_t *__restrict__ dlo, uint32_t *__restrict__ dhi)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t g = ex[i];
    gstart[g] = (uint32_t)i;
    gfirst[g] = P[i];   // stable sort: the smallest flat index of the group
    dlo[g] = hlo[P[i]];
    dhi[g] = shi[i];
}

__global__ void k_group_df(const uint32_t *__restrict__ gstart, int64_t nd, int64_t n, uint32_t *__restrict__ df)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nd) df[g] = (g + 1 < nd ? gstart[g + 1] : (uint32_t)n) - gstart[g];
}

__global__ void k_rank(const uint32_t *__restrict__ order, int64_t nd, uint32_t *__restrict__ rank)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nd) rank[order[k]] = (uint32_t)k;
}

__global__ void k_tid(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ex,
                      const uint32_t *__restrict__ P, const uint32_t *__restrict__ rank, int64_t n,
                      uint32_t *__restrict__ tid)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tid[P[i]] = rank[seg_of(flag, ex, i)];
}

__global__ void k_nonzero(const uint32_t *__restrict__ v, int64_t n, uint32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = v[i] ? 1u : 0u;
}

// lines with q-grams: their line index and q-gram count, in line order
__global__ void k_rec_compact(const uint32_t *__restrict__ ng, const uint32_t *__restrict__ off, int64_t nl,
                              uint32_t *__restrict__ rline, uint32_t *__restrict__ rlen)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nl || ng[r] == 0) return;
    rline[off[r]] = (uint32_t)r;
    rlen[off[r]] = ng[r];
}

// sorted position k (ascending count, stable) -> internal index n - 1 - k
__global__ void k_rec_order(const uint32_t *__restrict__ sline, const uint32_t *__restrict__ ng, int64_t n,
                            uint32_t *__restrict__ intern, uint32_t *__restrict__ id, uint32_t *__restrict__ len)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t r = sline[k], x = (uint32_t)(n - 1 - k);
    intern[r] = x;
    id[x] = r + 1u;
    len[x] = ng[r];
}

__global__ void k_out(const uint32_t *__restrict__ P, const uint32_t *__restrict__ tid, const uint32_t *__restrict__ gpos,
                      int64_t n, uint32_t *__restrict__ tok, uint32_t *__restrict__ loc)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    tok[k] = tid[P[k]];
    loc[k] = gpos[P[k]];
}

inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>((n + t - 1) / t, 1); }

int bits_for(uint64_t maxv)
{
    int b = 1;
    while (b < 64 && (1ull << b) <= maxv) b++;
    return b;
}

// device allocations released together
struct DevArena {
    std::vector<void *> ptrs;
    ~DevArena()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    uint32_t *u32(int64_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, (size_t)std::max<int64_t>(n, 1) * 4) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (uint32_t *)p;
    }
};

struct StreamGuard {
    hipStream_t st = nullptr;
    ~StreamGuard()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
};

}  // namespace

// Tokenize text on the current device; see the header comment for the steps.
extern "C" int fgss_collection_from_text_device(const char *text, int64_t nbytes, int q, fgss_collection **out)
{
    if (!out || (!text && nbytes > 0) || nbytes < 0) return fgss_fail(FGSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (q < 1 || q > 1024) return fgss_fail(FGSS_ERR_ARG, "q must be in 1..1024");   // preprocess_cli.cpp:86
    if (q > 8) return fgss_fail(FGSS_ERR_UNSUPPORTED, "the device tokenizer packs q-grams of at most 8 bytes");
    if (nbytes >= (int64_t)UINT32_MAX) return fgss_fail(FGSS_ERR_UNSUPPORTED, "text of 4 GiB or more");
    StreamGuard sg;
    TKCK(hipStreamCreateWithFlags(&sg.st, hipStreamNonBlocking));
    hipStream_t st = sg.st;
    DevArena A;
    auto *c = new fgss_collection();
    std::unique_ptr<fgss_collection> own(c);
    c->q = q;
    uint32_t *totals = A.u32(8);
    uint32_t *h_tot = nullptr;
    TKCK(hipHostMalloc((void **)&h_tot, 8 * 4, hipHostMallocDefault));
    struct PinnedFree {
        uint32_t *p;
        ~PinnedFree() { (void)hipHostFree(p); }
    } pf{h_tot};
    if (!totals) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    auto read_total = [&](int k) -> int {
        TKCK(hipMemcpyAsync(h_tot + k, totals + k, 4, hipMemcpyDeviceToHost, st));
        TKCK(hipStreamSynchronize(st));
        return FGSS_OK;
    };

    // 1. newlines
    uint8_t *dtext = (uint8_t *)A.u32((nbytes + 3) / 4 + 4);
    const int64_t ntiles = std::max<int64_t>((nbytes + kTileBytes - 1) / kTileBytes, 1);
    uint32_t *tcnt = A.u32(ntiles), *toff = A.u32(ntiles + 1);
    uint32_t *scr_small = A.u32(scan_scratch_elems(ntiles + 1));
    if (!dtext || !tcnt || !toff || !scr_small) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    if (nbytes) TKCK(hipMemcpyAsync(dtext, text, (size_t)nbytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_nl_count, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st, dtext, nbytes, tcnt);
    TKCK(hipGetLastError());
    TKCK(exclusive_scan<uint32_t>(tcnt, toff, ntiles, totals + 0, scr_small, st));
    int rc = read_total(0);
    if (rc) return rc;
    const int64_t n_nl = h_tot[0], nlines = n_nl + 1;
    uint32_t *nlpos = A.u32(n_nl);
    uint32_t *lstart = A.u32(nlines), *llen = A.u32(nlines), *lflag = A.u32(nlines), *loff = A.u32(nlines + 1);
    uint32_t *scr_lines = A.u32(scan_scratch_elems(nlines + 1));
    if (!nlpos || !lstart || !llen || !lflag || !loff || !scr_lines)
        return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    hipLaunchKernelGGL(k_nl_write, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st, dtext, nbytes, toff, nlpos);
    TKCK(hipGetLastError());

    // 2. non-empty lines and their q-gram counts
    hipLaunchKernelGGL(k_lines, dim3(blocks_for(nlines)), dim3(256), 0, st, nlpos, n_nl, nbytes, lstart, llen, lflag);
    TKCK(hipGetLastError());
    TKCK(exclusive_scan<uint32_t>(lflag, loff, nlines, totals + 1, scr_lines, st));
    rc = read_total(1);
    if (rc) return rc;
    const int64_t nl = h_tot[1];
    uint32_t *ls = A.u32(nl), *ll = A.u32(nl), *ng = A.u32(nl), *roff = A.u32(nl + 1);
    if (!ls || !ll || !ng || !roff) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    hipLaunchKernelGGL(k_lines_compact, dim3(blocks_for(nlines)), dim3(256), 0, st, lstart, llen, lflag, loff, nlines, q,
                       ls, ll, ng);
    TKCK(hipGetLastError());
    TKCK(exclusive_scan<uint32_t>(ng, roff, nl, totals + 2, scr_lines, st));
    rc = read_total(2);
    if (rc) return rc;
    const int64_t G = h_tot[2];

    // the per-q-gram arrays
    uint32_t *klo = A.u32(G), *khi = q > 4 ? A.u32(G) : nullptr, *gline = A.u32(G), *gpos = A.u32(G);
    uint32_t *occ = A.u32(G), *flag = A.u32(G), *ex = A.u32(G + 1), *aux = A.u32(G), *aux2 = A.u32(G);
    uint32_t *ka = A.u32(G), *va = A.u32(G), *kb = A.u32(G), *vb = A.u32(G);
    uint32_t *sscr = A.u32(std::max(radix_scratch_elems(G), scan_scratch_elems(G + 1)));
    if (!klo || (q > 4 && !khi) || !gline || !gpos || !occ || !flag || !ex || !aux || !aux2 || !ka || !va || !kb ||
        !vb || !sscr)
        return fgss_fail(FGSS_ERR_NOMEM, "device allocation for %lld q-grams failed", (long long)G);
    // stable sort of (keys, vals) copied into (ka, va): vals null = identity; the
    // sorted keys and values are returned through *ks, *vs
    auto sort = [&](const uint32_t *keys, const uint32_t *vals, int64_t n, int bits, uint32_t **ks,
                    uint32_t **vs) -> int {
        if (n <= 0) { *ks = ka; *vs = va; return FGSS_OK; }
        if (keys != ka) TKCK(hipMemcpyAsync(ka, keys, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        if (vals && vals != va) TKCK(hipMemcpyAsync(va, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        int which = 0;
        TKCK(radix_sort_kv(ka, va, kb, vb, n, nullptr, bits, vals != nullptr, sscr, st, &which));
        *ks = which ? kb : ka;
        *vs = which ? vb : va;
        return FGSS_OK;
    };
    uint32_t *ks = nullptr, *vs = nullptr;

    if (G > 0) {
        // 3. keys
        hipLaunchKernelGGL(k_grams, dim3((unsigned)std::min<int64_t>(blocks_for(nl * kWave), 16384)), dim3(256), 0, st,
                           dtext, ls, ng, roff, nl, q, klo, khi, gline, gpos);
        TKCK(hipGetLastError());
        // 4. occurrence numbers
        rc = sort(klo, nullptr, G, 8 * std::min(q, 4), &ks, &vs);
        if (rc) return rc;
        if (q > 4) {
            hipLaunchKernelGGL(k_gather, dim3(blocks_for(G)), dim3(256), 0, st, aux, khi, vs, G);
            TKCK(hipGetLastError());
            TKCK(hipMemcpyAsync(aux2, vs, (size_t)G * 4, hipMemcpyDeviceToDevice, st));
            rc = sort(aux, aux2, G, 8 * (q - 4), &ks, &vs);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_occ_flags, dim3(blocks_for(G)), dim3(256), 0, st, vs, klo, khi, gline, G, flag);
        TKCK(hipGetLastError());
        TKCK(exclusive_scan<uint32_t>(flag, ex, G, totals + 3, sscr, st));
        hipLaunchKernelGGL(k_seg_start, dim3(blocks_for(G)), dim3(256), 0, st, flag, ex, G, aux);
        TKCK(hipGetLastError());
        hipLaunchKernelGGL(k_occ, dim3(blocks_for(G)), dim3(256), 0, st, flag, ex, aux, vs, G, occ);
        TKCK(hipGetLastError());
        // 5. hashes (klo, khi are free again: hash low and high words)
        uint32_t *hlo = klo, *hhi = aux2;
        hipLaunchKernelGGL(k_hash, dim3(blocks_for(G)), dim3(256), 0, st, dtext, ls, gline, gpos, occ, G, q, hlo, hhi);
        TKCK(hipGetLastError());
        // 6. stable sort by the 64-bit hash: low word, then high word
        rc = sort(hlo, nullptr, G, 32, &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gather, dim3(blocks_for(G)), dim3(256), 0, st, aux, hhi, vs, G);
        TKCK(hipGetLastError());
        TKCK(hipMemcpyAsync(occ, vs, (size_t)G * 4, hipMemcpyDeviceToDevice, st));   // occ is done with
        rc = sort(aux, occ, G, 32, &ks, &vs);
        if (rc) return rc;
        uint32_t *P = occ, *shi = aux;   // keep the final order and high words outside the sort buffers
        TKCK(hipMemcpyAsync(P, vs, (size_t)G * 4, hipMemcpyDeviceToDevice, st));
        TKCK(hipMemcpyAsync(shi, ks, (size_t)G * 4, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_hash_flags, dim3(blocks_for(G)), dim3(256), 0, st, P, hlo, shi, G, flag);
        TKCK(hipGetLastError());
        TKCK(exclusive_scan<uint32_t>(flag, ex, G, totals + 4, sscr, st));
        rc = read_total(4);
        if (rc) return rc;
    }
    const int64_t nd = G > 0 ? h_tot[4] : 0;
    uint32_t *gstart = A.u32(nd), *gfirst = A.u32(nd), *dlo = A.u32(nd), *dhi = A.u32(nd), *df = A.u32(nd),
             *rank = A.u32(nd);
    if (!gstart || !gfirst || !dlo || !dhi || !df || !rank) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    uint32_t *tid = aux2;   // the hash high words are consumed by then
    if (G > 0) {
        uint32_t *P = occ, *shi = aux;
        hipLaunchKernelGGL(k_group_info, dim3(blocks_for(G)), dim3(256), 0, st, flag, ex, P, klo, shi, G, gstart, gfirst,
                           dlo, dhi);
        TKCK(hipGetLastError());
        hipLaunchKernelGGL(k_group_df, dim3(blocks_for(nd)), dim3(256), 0, st, gstart, nd, G, df);
        TKCK(hipGetLastError());
        // 7. token id = rank by (df, first occurrence)
        const int gb = bits_for((uint64_t)G);
        rc = sort(gfirst, nullptr, nd, gb, &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gather, dim3(blocks_for(nd)), dim3(256), 0, st, gstart, df, vs, nd);   // gstart is done with
        TKCK(hipGetLastError());
        rc = sort(gstart, vs, nd, gb, &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rank, dim3(blocks_for(nd)), dim3(256), 0, st, vs, nd, rank);
        TKCK(hipGetLastError());
        hipLaunchKernelGGL(k_tid, dim3(blocks_for(G)), dim3(256), 0, st, flag, ex, P, rank, G, tid);
        TKCK(hipGetLastError());
    }

    // 8. records: lines with q-grams, stable by count; then tokens within records
    uint32_t *rflag = A.u32(nl), *roff2 = A.u32(nl + 1), *rline = A.u32(nl), *rlen = A.u32(nl);
    if (!rflag || !roff2 || !rline || !rlen) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    hipLaunchKernelGGL(k_nonzero, dim3(blocks_for(nl)), dim3(256), 0, st, ng, nl, rflag);
    TKCK(hipGetLastError());
    TKCK(exclusive_scan<uint32_t>(rflag, roff2, nl, totals + 5, scr_lines, st));
    rc = read_total(5);
    if (rc) return rc;
    const int64_t n = nl > 0 ? h_tot[5] : 0;
    uint32_t *intern = A.u32(nl), *did = A.u32(n), *dlen = A.u32(n), *dtok = A.u32(G), *dloc = A.u32(G);
    if (!intern || !did || !dlen || !dtok || !dloc) return fgss_fail(FGSS_ERR_NOMEM, "device allocation failed");
    if (n > 0) {
        hipLaunchKernelGGL(k_rec_compact, dim3(blocks_for(nl)), dim3(256), 0, st, ng, roff2, nl, rline, rlen);
        TKCK(hipGetLastError());
        rc = sort(rlen, rline, n, bits_for((uint64_t)G), &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rec_order, dim3(blocks_for(n)), dim3(256), 0, st, vs, ng, n, intern, did, dlen);
        TKCK(hipGetLastError());
        // per q-gram: its record's internal index; sort by token id, then by record
        rc = sort(tid, nullptr, G, bits_for((uint64_t)std::max<int64_t>(nd, 1)), &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gather, dim3(blocks_for(G)), dim3(256), 0, st, flag, gline, vs, G);   // line per sorted gram
        TKCK(hipGetLastError());
        hipLaunchKernelGGL(k_gather, dim3(blocks_for(G)), dim3(256), 0, st, ex, intern, flag, G);  // its record
        TKCK(hipGetLastError());
        rc = sort(ex, vs, G, bits_for((uint64_t)n), &ks, &vs);
        if (rc) return rc;
        hipLaunchKernelGGL(k_out, dim3(blocks_for(G)), dim3(256), 0, st, vs, tid, gpos, G, dtok, dloc);
        TKCK(hipGetLastError());
    }

    // results to the host collection
    c->id.resize((size_t)n);
    c->len.resize((size_t)n);
    if (n) {
        TKCK(hipMemcpyAsync(c->id.data(), did, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        TKCK(hipMemcpyAsync(c->len.data(), dlen, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    }
    TKCK(hipStreamSynchronize(st));
    (void)fgss_finish_collection(c);   // offsets and lengths (no tokens yet; n_terms is set below)
    c->tok.resize((size_t)G);
    c->loc.resize((size_t)G);
    c->dict_hash.resize((size_t)nd);
    c->dict_id.resize((size_t)nd);
    std::vector<uint32_t> hlo_d((size_t)nd), hhi_d((size_t)nd);
    if (G) {
        TKCK(hipMemcpyAsync(c->tok.data(), dtok, (size_t)G * 4, hipMemcpyDeviceToHost, st));
        TKCK(hipMemcpyAsync(c->loc.data(), dloc, (size_t)G * 4, hipMemcpyDeviceToHost, st));
    }
    if (nd) {
        TKCK(hipMemcpyAsync(hlo_d.data(), dlo, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
        TKCK(hipMemcpyAsync(hhi_d.data(), dhi, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
        TKCK(hipMemcpyAsync(c->dict_id.data(), rank, (size_t)nd * 4, hipMemcpyDeviceToHost, st));
    }
    TKCK(hipStreamSynchronize(st));
    for (int64_t g = 0; g < nd; g++)
        c->dict_hash[(size_t)g] = ((uint64_t)hhi_d[(size_t)g] << 32) | hlo_d[(size_t)g];
    if (nd >= (int64_t)INT32_MAX) return fgss_fail(FGSS_ERR_UNSUPPORTED, "too many distinct q-grams");
    c->n_terms = (int32_t)nd;
    *out = own.release();
    return FGSS_OK;
}

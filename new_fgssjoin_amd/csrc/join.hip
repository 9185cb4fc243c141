// join.hip -- the fgssjoin self-join on MI355X (gfx950).
//
// Pipeline per threshold t (all device work on one HIP stream):
//   1. k_midlen        mid_prefix(|s|) per record          (src/fgssjoin/data.cu:241-257)
//   2. exclusive_scan   -> entry offsets                     (data.cu:181, scan.cuh:39-105)
//   3. k_entries       (term, entry) for every mid-prefix token (index.cu:276-296)
//   4. radix_sort_pairs stable by term -> postings ascending by record id
//                        (replaces df_count_kernel + inverted_index_kernel, index.cu:303-342,
//                         whose list order was arbitrary)
//   5. k_postings      SoA postings (sid, pos) + per-term list offsets
//   6. k_lmin          per length l: the smallest partner length passing the size filter
//   7. k_plan          per query: source window (q, hi] and the number of useful postings
//   8. k_items         split heavy queries into source sub-windows ("work items")
//   9. k_join          one wavefront per work item: candidate generation with overlap
//                      accumulation in an LDS hash table (filtering.cu:76-141), positional
//                      filter, verification against an LDS bitmap of the query tokens with
//                      ballot/popcount (checking.cu:93-175), float32 final test
//                      (block.cu:183-197) and compaction of result pairs.
//
// Result semantics (identical to the reference's printed pair set): a pair (q, s),
// q < s in internal order, is a result iff
//   size filter holds, m = |Q[0,qp) n S[0,sp)| > 0, the positional filter passes,
//   ov = |Q n S| >= floor(bound) and (float)ov/(float)(|Q|+|S|-ov) >= t.
// The positional filter depends only on (m, last matched query position, last matched
// source position) because both records share one global token order, so it can be
// accumulated in any order with an atomic add and an atomic max.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "fgss_bounds.h"
#include "fgss_internal.h"
#include "prims.h"

using namespace fgss;

#define HIPCK(x)                                                                                    \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fgss_fail(FGSS_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// wavefronts per block in k_join: four, or two for the largest LDS tables
constexpr int join_waves(int cap) { return cap >= 4096 ? 2 : 4; }
constexpr int kTokBatch = 128;             // probe tokens whose posting ranges are staged at once
constexpr int kQueueBatch = 16;            // work items claimed per queue atomic
constexpr int kJobCtr = 4;                 // per-job counters: postings, candidates, intersections, results
constexpr uint32_t kFence = 32;            // postings per fence block (one 128-byte line of ids)
constexpr int kSigBuckets = 128;           // token buckets of a record signature
constexpr int kSigWords = kSigBuckets / 4; // one byte-wide count per bucket
constexpr int kSig64Words = 16;            // the same counts merged pairwise into 64 buckets

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// device counters, one 64-bit word each
enum Ctr {
    C_PQ = 0,
    C_ESCAN,
    C_CAND,
    C_INTER,
    C_VTOK,
    C_SIGCHK,        // candidates tested against the signature bound
    C_VERIFIED,      // candidates whose tokens were merged
    C_SIG2,          // candidates that needed the 128-bucket signature
    C_RESULTS,       // pairs emitted (may exceed the output capacity)
    C_OVERFLOW,      // work items whose LDS table overflowed
    C_QUEUE,         // work-queue head
    C_QUEUE_D,       // work-queue head of the dense role (dense units)
    C_CYC_SETUP,     // shader-clock cycles summed over waves, per k_join phase
    C_CYC_A,
    C_CYC_B,
    C_CYC_C,
    C_CYC_D,
    C_CYC_RANGES,    // part of C_CYC_A spent locating posting ranges
    C_RETRY,         // streamed windows of several parts that overflowed and ran again part by part
    C_DENSE,         // dense parts joined
    C_DPOST,         // postings scanned by dense parts (4-byte postings)
    C_COUNT
};

struct Item {
    uint32_t q, a, b;   // query and source sub-window [a, b)
    uint32_t pad;       // expected useful postings (sizes the LDS table) | kWholeWindow | kPlanned | kStreamed
    uint32_t part, nparts;  // planned items: part index and part count of the query
    uint32_t bofs;          // planned items: first boundary position of the query in the bounds table
    uint32_t job;           // the query's job
    float t;                // its threshold
    uint32_t lq, qpos, qpoff;   // the query's length (| kMidProbe), token offset and probe-range offset
};
constexpr int kItemWords = (int)(sizeof(Item) / 4);
static_assert(sizeof(Item) == 48 && offsetof(Item, q) == 0 && offsetof(Item, a) == 4 && offsetof(Item, b) == 8 &&
                  offsetof(Item, pad) == 12 && offsetof(Item, part) == 16 && offsetof(Item, nparts) == 20 &&
                  offsetof(Item, bofs) == 24 && offsetof(Item, job) == 28 && offsetof(Item, t) == 32 &&
                  offsetof(Item, lq) == 36 && offsetof(Item, qpos) == 40 && offsetof(Item, qpoff) == 44,
              "k_join reads item fields by word index");

// word `lane` of item i for lanes below kItemWords (the others read word 0)
__device__ __forceinline__ uint32_t item_word(const Item *__restrict__ items, uint32_t i, int lane)
{
    return ((const uint32_t *)(items + i))[lane < kItemWords ? lane : 0];
}

// R-S items over the sources longer than the probe: the probe contributes its mid
// prefix and the sources their max prefix (the longer record is the query)
constexpr uint32_t kMidProbe = 0x10000u;
// an item of a job whose lengths are not non-increasing: every posting of the window
// passes the 4-way size filter itself (a sorted job's window holds only partners)
constexpr uint32_t kSizeFilter = 0x20000u;
// jend[job] | kJobUnsorted: the job's record lengths are not non-increasing (its
// queries' windows then run to the job's end)
constexpr uint32_t kJobUnsorted = 0x80000000u;
constexpr uint32_t kWholeWindow = 0x80000000u;   // the item covers the query's whole window
constexpr uint32_t kPlanned = 0x40000000u;       // part of a planned split: ranges come from the bounds table
constexpr uint32_t kStreamed = 0x20000000u;      // the whole window, walked part by part by one wavefront
constexpr uint32_t kPadMask = 0x0FFFFFFFu;
// nparts flag of a query joined in dense parts: the parts of a global grid of 2 x
// table-slots records that meet its window, each counted in a table indexed by the
// source (no hashing); the count is that of its dense parts
constexpr uint32_t kDenseParts = 0x80000000u;
// streamed items keep one posting cursor per probe token in the wavefront's LDS
// (the range arrays of the other items)
constexpr uint32_t kStreamQp = 256;
// default: only queries whose boundary rows do not fit the table are streamed (for
// short probes the planned parts spread better over the wavefronts)
constexpr uint32_t kStreamParts = 0xFFFFFFFFu;
constexpr uint32_t kHeavyParts = 64;   // queries split into more parts use the heavy target
#ifndef FGSS_STREAM_CHUNK
#define FGSS_STREAM_CHUNK 64
#endif
constexpr uint32_t kStreamChunk = FGSS_STREAM_CHUNK;   // parts per streamed item (a heavy query gets several)
#ifndef FGSS_MAX_SPAN
#define FGSS_MAX_SPAN 16
#endif
constexpr uint32_t kMaxSpan = FGSS_MAX_SPAN;     // parts per window of a streamed item at most
#ifndef FGSS_SPAN_GROW_PCT
#define FGSS_SPAN_GROW_PCT 40
#endif
#ifndef FGSS_SPAN_SHRINK_PCT
#define FGSS_SPAN_SHRINK_PCT 70
#endif
// table fill (% of its slots) after which a streamed item's window span doubles / halves
constexpr uint32_t kSpanGrowPct = FGSS_SPAN_GROW_PCT, kSpanShrinkPct = FGSS_SPAN_SHRINK_PCT;

// How the items of a query with np parts read their posting ranges.
enum ItemMode : uint32_t { IM_NONE = 0, IM_WHOLE, IM_PLANNED, IM_STREAMED, IM_SEARCHED, IM_DENSE };
__device__ __forceinline__ uint32_t item_mode(uint32_t np, uint32_t qp, uint64_t bcnt, uint64_t bofs, uint32_t cap_bounds)
{
    if (np & kDenseParts) return IM_DENSE;
    if (np == 0) return IM_NONE;
    if (np == 1) return IM_WHOLE;
    if (bcnt > 0 && bofs + bcnt <= cap_bounds) return IM_PLANNED;
    return qp <= kStreamQp ? IM_STREAMED : IM_SEARCHED;
}

}  // namespace

// Records resident on one device: one collection, or the concatenation of the
// collections of a batch (each record carries the id of its job; a job's records
// are contiguous and end at jend[job]).
struct fgss_device_sets {
    int device = 0;
    hipStream_t st = nullptr;
    int64_t n = 0, ntok = 0;
    uint32_t n_terms = 0, njobs = 1;
    int32_t max_len = 0;
    uint32_t *tok = nullptr, *len = nullptr, *pos = nullptr;
    uint32_t *sig = nullptr;    // kSigWords words per record (token-bucket byte counts)
    uint32_t *sig64 = nullptr;  // kSig64Words words per record (64-bucket byte counts)
    uint8_t *sigbad = nullptr;  // per record: a count did not fit in a byte
    uint32_t *job = nullptr;    // per record: job id
    uint32_t *jend = nullptr;   // per job: one past its last record
    float *tj = nullptr;        // per job: threshold of the current run
    DevBuf b_psid, b_fence, b_bcnt, b_bofs, b_bounds, b_probelen, b_poff, b_rng, b_eq, b_midlen, b_midoff, b_ka,
        b_ia, b_kb, b_ib, b_sortscr, b_sid_of, b_pl_of, b_post, b_loff, b_hi, b_nparts, b_partoff, b_scanscr, b_items,
        b_items2, b_ovf, b_out, b_ctr, b_jctr, b_misc, b_nitems, b_ovfp, b_wctr, b_jsave, b_dlist, b_dflag, b_doff, b_grp, b_pden,
        b_doffs, b_dq, b_dtpre;
    const uint32_t *skeys = nullptr;   // the last index build's sorted entry terms, one per posting
    uint32_t cap_items = 0, cap_bounds = 0, cap_out = 0;   // current capacities (elements)
    uint32_t cap_ovf = 0;         // overflow records (a streamed item can record one per part)
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint64_t *h_ctr = nullptr;    // pinned mirror of the counters
    uint32_t *h_misc = nullptr;   // pinned mirror of the device totals
    uint64_t *h_jctr = nullptr;   // pinned mirror of the per-job counters
    bool jctr_async = false;      // the run's tail copies the per-job counters into h_jctr
    int32_t *h_out = nullptr;     // pinned mirror of the first out_copied result triples
    size_t h_out_cap = 0;         // triples h_out holds
    uint32_t out_hint = 0;        // triples the run's tail copies: the most a run has found, plus a margin
    uint32_t out_copied = 0;      // triples the last run's tail copied
    // captured pipeline stages (index build, planning, join kernel), replayed with the
    // timing events and the counter copies enqueued between and after them
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
    std::array<int64_t, 17> graph_key{};   // what the captured stages were built for (graph_key_of)
    int cap_gen = 0;              // bumped whenever a buffer the pipeline uses is reallocated
    bool sized = false;           // a run has completed (first-run buffer sizing is over)
    int32_t q_begin = 0, q_end = 0;   // the run's query range (internal indices; 0, 0: every record)
    // a batch run's job group: only records [grp_begin, grp_end) are indexed and probe
    // (0, 0: every record); written to b_grp before each run, outside any captured graph
    uint32_t grp_begin = 0, grp_end = 0;
    // Buffers replaced by larger ones (device, pinned host): freed when the record set is
    // released.  hipFree and hipHostFree wait for the whole device, which would stall the
    // runs of other record sets (a server's other slots, another thread's collection).
    std::vector<void *> dead_dev, dead_host;
};

static int ensure(fgss_device_sets *d, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return FGSS_OK;
    // a buffer that grows again grows by half at least, so that a workspace that creeps
    // upward (a server's slots) reallocates rarely and its retired buffers stay few
    size_t want = std::max(bytes + bytes / 8, b.cap + b.cap / 2);
    if (b.p) d->dead_dev.push_back(b.p);   // work already enqueued on the stream may still read it
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) return fgss_fail(FGSS_ERR_NOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    b.cap = want;
    return FGSS_OK;
}

// A pinned host buffer replaced by a larger one: freed with the record set.
static void retire_host(fgss_device_sets *d, void *p)
{
    if (p) d->dead_host.push_back(p);
}

template <class T>
static T *P(DevBuf &b) { return (T *)b.p; }


void fgss_device_release(fgss_collection *c)
{
    if (!c || !c->dev) return;
    fgss_device_sets *d = c->dev;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(d->device);
    if (d->st) (void)hipStreamSynchronize(d->st);
    for (int k = 0; k < 3; k++) {
        if (d->graph_exec[k]) (void)hipGraphExecDestroy(d->graph_exec[k]);
        if (d->graph[k]) (void)hipGraphDestroy(d->graph[k]);
    }
    DevBuf *bufs[] = {&d->b_psid, &d->b_fence, &d->b_bcnt, &d->b_bofs, &d->b_bounds, &d->b_probelen, &d->b_poff,
                      &d->b_rng, &d->b_eq, &d->b_midlen, &d->b_midoff, &d->b_ka, &d->b_ia, &d->b_kb, &d->b_ib,
                      &d->b_sortscr, &d->b_sid_of, &d->b_pl_of, &d->b_post, &d->b_loff, &d->b_hi, &d->b_nparts, &d->b_partoff,
                      &d->b_scanscr, &d->b_items, &d->b_items2, &d->b_ovf, &d->b_out, &d->b_ctr, &d->b_jctr,
                      &d->b_nitems, &d->b_ovfp, &d->b_wctr, &d->b_jsave, &d->b_dlist, &d->b_dflag, &d->b_doff, &d->b_grp, &d->b_pden,
                      &d->b_doffs, &d->b_dq, &d->b_dtpre, &d->b_misc};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    void *arrs[] = {d->tok, d->len, d->pos, d->sig, d->sig64, d->sigbad, d->job, d->jend, d->tj};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    for (void *a : d->dead_dev) (void)hipFree(a);
    for (void *a : d->dead_host) (void)hipHostFree(a);
    for (auto &e : d->ev)
        if (e) (void)hipEventDestroy(e);
    if (d->h_ctr) (void)hipHostFree(d->h_ctr);
    if (d->h_misc) (void)hipHostFree(d->h_misc);
    if (d->h_jctr) (void)hipHostFree(d->h_jctr);
    if (d->h_out) (void)hipHostFree(d->h_out);
    if (d->st) (void)hipStreamDestroy(d->st);
    (void)hipSetDevice(cur);
    delete d;
    c->dev = nullptr;
}

// ---------------------------------------------------------------------------
// record signatures: per record, how many of its tokens fall in each of 128
// hash buckets, as bytes.  For any two records sum_b min(cQ[b], cS[b]) bounds
// their overlap from above, and equals (|Q| + |S| - sum_b |cQ[b] - cS[b]|) / 2,
// which v_sad_u8 evaluates four buckets per instruction.  Built once per upload.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sig_bucket(uint32_t tok) { return (tok * 0x9E3779B1u) >> (32 - 7); }
// zero bytes past a collection's signature-overflow flags: a whole grid part of dense units
// (at most 2 x 4096 records) can be read past the last record
constexpr size_t kSigbadPad = 8192;

__global__ __launch_bounds__(256) void k_signature(const uint32_t *__restrict__ tok, const uint32_t *__restrict__ pos,
                                                   const uint32_t *__restrict__ len, int64_t n,
                                                   uint32_t *__restrict__ sig, uint32_t *__restrict__ sig64,
                                                   uint8_t *__restrict__ sigbad)
{
    __shared__ uint32_t hist[256 / kWave][kSigBuckets];
    const int w = (int)threadIdx.x / kWave;
    const int lane = lane_id();
    uint32_t *h = hist[w];
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + w; r < n; r += nwaves) {
        for (int b = lane; b < kSigBuckets; b += kWave) h[b] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t l = len[r], p = pos[r];
        for (uint32_t i = lane; i < l; i += kWave) atomicAdd(&h[sig_bucket(tok[p + i])], 1u);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool bad = false;
        if (lane < kSigWords) {
            uint32_t c0 = h[4 * lane], c1 = h[4 * lane + 1], c2 = h[4 * lane + 2], c3 = h[4 * lane + 3];
            sig[(size_t)r * kSigWords + lane] = (c0 & 255u) | (c1 & 255u) << 8 | (c2 & 255u) << 16 | (c3 & 255u) << 24;
        }
        if (lane < kSig64Words) {   // bucket b of 64 = buckets 2b, 2b+1 of 128; every count <= its merge
            uint32_t m[4];
#pragma unroll
            for (int k = 0; k < 4; k++) m[k] = h[8 * lane + 2 * k] + h[8 * lane + 2 * k + 1];
            bad = (m[0] | m[1] | m[2] | m[3]) > 255u;
            sig64[(size_t)r * kSig64Words + lane] = m[0] | m[1] << 8 | m[2] << 16 | m[3] << 24;
        }
        const uint64_t anybad = __ballot(bad);
        if (lane == 0) sigbad[r] = anybad ? 1 : 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// upload (the H2D part of prepare_data, src/fgssjoin/data.cu:133-153)
// ---------------------------------------------------------------------------
static int validate_collection(const fgss_collection *c)
{
    const int64_t n = (int64_t)c->len.size();
    if (n >= (int64_t)INT32_MAX) return fgss_fail(FGSS_ERR_UNSUPPORTED, "more than 2^31-1 records");
    if ((int64_t)c->tok.size() >= (int64_t)UINT32_MAX)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "more than 2^32-1 tokens (the reference's unsigned offsets)");
    if (c->max_len >= 65536) return fgss_fail(FGSS_ERR_UNSUPPORTED, "records longer than 65535 tokens");
    if (c->n_terms < 0 || c->n_terms >= INT32_MAX)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "token ids must be below 2^31-1");
    for (int64_t i = 0; i < n; i++) {
        const int32_t *r = c->tok.data() + c->pos[i];
        for (int32_t k = 0; k < c->len[i]; k++) {
            if (r[k] < 0) return fgss_fail(FGSS_ERR_UNSUPPORTED, "negative token id in record %lld", (long long)i);
            // the index's list offsets have n_terms + 1 entries: a larger id would index past them
            if (r[k] >= c->n_terms)
                return fgss_fail(FGSS_ERR_UNSUPPORTED, "record %lld: token id %d not below n_terms %d", (long long)i,
                                 r[k], c->n_terms);
            if (k > 0 && r[k] <= r[k - 1])
                return fgss_fail(FGSS_ERR_UNSUPPORTED, "record %lld: tokens must be strictly ascending", (long long)i);
        }
    }
    return FGSS_OK;
}

// Upload a collection into a new device record set (validated first).
static int upload_records(fgss_device_sets **out, const fgss_collection *c)
{
    *out = nullptr;
    int rc = validate_collection(c);
    if (rc) return rc;
    auto *d = new fgss_device_sets();
    fgss_collection holder;   // releases d on failure
    holder.dev = d;
    HIPCK(hipGetDevice(&d->device));
    HIPCK(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking));
    for (auto &e : d->ev) HIPCK(hipEventCreate(&e));
    d->n = (int64_t)c->len.size();
    d->ntok = (int64_t)c->tok.size();
    d->n_terms = (uint32_t)c->n_terms;
    d->max_len = c->max_len;
    d->njobs = c->jend.empty() ? 1u : (uint32_t)c->jend.size();
    HIPCK(hipHostMalloc((void **)&d->h_ctr, C_COUNT * sizeof(uint64_t), hipHostMallocDefault));
    HIPCK(hipHostMalloc((void **)&d->h_misc, 16 * sizeof(uint32_t), hipHostMallocDefault));
    HIPCK(hipHostMalloc((void **)&d->h_jctr, (size_t)d->njobs * kJobCtr * sizeof(uint64_t), hipHostMallocDefault));
    const size_t nb = (size_t)std::max<int64_t>(d->n, 1) * 4, tb = (size_t)std::max<int64_t>(d->ntok, 1) * 4;
    const size_t jb = (size_t)d->njobs * 4;
    if (hipMalloc(&d->tok, tb) != hipSuccess || hipMalloc(&d->len, nb) != hipSuccess ||
        hipMalloc(&d->pos, nb) != hipSuccess || hipMalloc(&d->sig, nb * kSigWords) != hipSuccess ||
        hipMalloc(&d->sig64, nb * kSig64Words) != hipSuccess || hipMalloc(&d->sigbad, nb + kSigbadPad) != hipSuccess ||
        hipMalloc(&d->job, nb) != hipSuccess || hipMalloc(&d->jend, jb) != hipSuccess ||
        hipMalloc(&d->tj, jb) != hipSuccess) {
        return fgss_fail(FGSS_ERR_NOMEM, "device allocation for the collection failed");
    }
    // each job's end, flagged when its lengths are not non-increasing in internal order
    // (a .bin in any record order: its queries' windows run to the job's end)
    std::vector<uint32_t> je(d->njobs);
    {
        int64_t start = 0;
        for (uint32_t j = 0; j < d->njobs; j++) {
            const int64_t end = c->jend.empty() ? d->n : (int64_t)c->jend[j];
            bool sorted = true;
            for (int64_t i = start + 1; i < end && sorted; i++) sorted = c->len[(size_t)i] <= c->len[(size_t)i - 1];
            je[j] = (uint32_t)end | (sorted ? 0u : kJobUnsorted);
            start = end;
        }
    }
    if (c->jend.empty()) HIPCK(fill<uint32_t>(d->job, d->n, 0u, d->st));
    else HIPCK(hipMemcpyAsync(d->job, c->job.data(), (size_t)d->n * 4, hipMemcpyHostToDevice, d->st));
    HIPCK(hipMemcpyAsync(d->jend, je.data(), jb, hipMemcpyHostToDevice, d->st));   // je lives until the sync below
    std::vector<uint32_t> pos32((size_t)d->n);
    for (int64_t i = 0; i < d->n; i++) pos32[(size_t)i] = (uint32_t)c->pos[(size_t)i];
    HIPCK(hipMemcpyAsync(d->tok, c->tok.data(), (size_t)d->ntok * 4, hipMemcpyHostToDevice, d->st));
    HIPCK(hipMemcpyAsync(d->len, c->len.data(), (size_t)d->n * 4, hipMemcpyHostToDevice, d->st));
    HIPCK(hipMemcpyAsync(d->pos, pos32.data(), (size_t)d->n * 4, hipMemcpyHostToDevice, d->st));
    // the overflow bytes past the last record read as zeros (dense units read whole grid parts)
    HIPCK(hipMemsetAsync(d->sigbad + d->n, 0, kSigbadPad, d->st));
    if (d->n > 0) {
        const int64_t blocks = std::min<int64_t>((d->n + 3) / 4, 65536);
        hipLaunchKernelGGL(k_signature, dim3((unsigned)blocks), dim3(256), 0, d->st, d->tok, d->pos, d->len, d->n,
                           d->sig, d->sig64, d->sigbad);
        HIPCK(hipGetLastError());
    }
    HIPCK(hipStreamSynchronize(d->st));
    holder.dev = nullptr;
    *out = d;
    return FGSS_OK;
}

extern "C" int fgss_collection_upload(fgss_collection *c)
{
    if (!c) return fgss_fail(FGSS_ERR_ARG, "null collection");
    if (c->dev) return FGSS_OK;
    return upload_records(&c->dev, c);
}

// ---------------------------------------------------------------------------
// index build kernels
// ---------------------------------------------------------------------------
// per record: mid-prefix length (indexed tokens) and probe-prefix length, both
// clamped to the record length so that an empty record contributes nothing
// With rs set the indexed prefix is the probe prefix max_prefix (an R-S index built
// at the smallest threshold it will serve) instead of the self-join mid-prefix.
// Records outside the query range [qb, qe) (unless qb = qe = 0: every record) probe with
// no tokens: they stay sources of the queries inside it, which are joined as usual.
// grp (null: none): a job group [grp[0], grp[1]) of a batch run (grp[1] = 0: every
// record); records outside it are neither indexed nor probe.
__global__ void k_midlen(const uint32_t *__restrict__ len, const uint32_t *__restrict__ job,
                         const float *__restrict__ tj, uint32_t *__restrict__ midlen,
                         uint32_t *__restrict__ probelen, int64_t n, int rs, int64_t qb, int64_t qe,
                         const uint32_t *__restrict__ grp)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float t = tj[job[i]];
        uint32_t l = len[i];
        uint32_t mp = rs ? fgss_max_prefix(l, t) : fgss_mid_prefix(l, t);
        uint32_t qp = fgss_max_prefix(l, t);
        const bool out = grp != nullptr && grp[1] != 0u && (i < (int64_t)grp[0] || i >= (int64_t)grp[1]);
        midlen[i] = out ? 0u : (mp < l ? mp : l);
        probelen[i] = out || ((qb != 0 || qe != 0) && (i < qb || i >= qe)) ? 0u : (qp < l ? qp : l);
    }
}

__global__ void k_set_pair(uint32_t *__restrict__ p, uint32_t a, uint32_t b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p[0] = a;
        p[1] = b;
    }
}

// Index entries in record order: for e in record r's indexed prefix, keys[e] = the
// token, sid_of[e] = r and pl_of[e] = its position | |r| << 16.  A block takes 256
// consecutive records, whose entries [midoff[r0], midoff[r0 + 256]) are contiguous,
// one entry per thread (coalesced stores); a thread finds its record among the
// block's by a search over their offsets in LDS.
__global__ __launch_bounds__(256) void k_entries(const uint32_t *__restrict__ tok, const uint32_t *__restrict__ pos,
                                                 const uint32_t *__restrict__ len,
                                                 const uint32_t *__restrict__ midoff,
                                                 const uint32_t *__restrict__ nE_dev, int64_t n,
                                                 uint32_t *__restrict__ keys, uint32_t *__restrict__ sid_of,
                                                 uint32_t *__restrict__ pl_of)
{
    __shared__ uint32_t soff[257], spos[256], slen[256];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    if (r0 >= n) return;
    const int nr = (int)min<int64_t>(256, n - r0);
    for (int i = threadIdx.x; i <= nr; i += 256) soff[i] = r0 + i < n ? midoff[r0 + i] : *nE_dev;
    for (int i = threadIdx.x; i < nr; i += 256) { spos[i] = pos[r0 + i]; slen[i] = len[r0 + i]; }
    __syncthreads();
    for (uint32_t e = soff[0] + threadIdx.x; e < soff[nr]; e += 256) {
        int lo = 0, hi = nr - 1;   // the last record of the block whose entries start at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (soff[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t j = e - soff[lo];
        keys[e] = tok[spos[lo] + j];
        sid_of[e] = (uint32_t)(r0 + lo);
        pl_of[e] = j | (slen[lo] << 16);
    }
}

// postings in term order, packed: x = record id, y = token position | record length << 16
// pden (null: none): the 4-byte postings that dense parts of dense_w records read,
// position << 24 | (record mod dense_w) << (24 - log2 dense_w) | record length (a posting
// a dense part reads has its position below 256, the longer ones are clamped and never
// read; dense parts are on only for lengths that fit their field)
__global__ void k_postings(const uint32_t *__restrict__ sidx, const uint32_t *__restrict__ sid_of,
                           const uint32_t *__restrict__ pl_of, const uint32_t *__restrict__ nE_dev,
                           uint2 *__restrict__ post, uint32_t *__restrict__ psid, uint32_t *__restrict__ fence,
                           uint32_t *__restrict__ pden, uint32_t dense_w, uint32_t dshift)
{
    const uint32_t nE = *nE_dev;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const uint32_t e = sidx[i];
    const uint32_t s = sid_of[e], pl = pl_of[e];
    post[i] = make_uint2(s, pl);
    psid[i] = s;
    if (pden) pden[i] = min(pl & 0xFFFFu, 255u) << 24 | (s & (dense_w - 1u)) << dshift | pl >> 16;
    if ((i & (kFence - 1)) == 0) fence[i / kFence] = s;
}

// loff[t] = first posting of term t; loff[n_terms] = nE
__global__ void k_list_offsets(const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ nE_dev,
                               uint32_t n_terms, uint32_t *__restrict__ loff)
{
    const uint32_t nE = *nE_dev;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nE) return;
    int64_t cur = i < nE ? (int64_t)skeys[i] : (int64_t)n_terms;
    int64_t prev = i > 0 ? (int64_t)skeys[i - 1] : -1;
    for (int64_t tt = prev + 1; tt <= cur; tt++) loff[tt] = i;
}

// Dense parts follow a global grid of dense_w records: grid part g holds the records
// [g * dense_w, (g + 1) * dense_w).  The offset table gives, for every term t and grid part
// g <= ng, the first posting of t's list whose record lies at or past g * dense_w:
// offs[g * n_terms + t] (the list's end when there is none), so that the postings of t in
// grid part g are offs[g][t] .. offs[g + 1][t] with no search.  The table holds the two
// together: entry (g, t) = (offs[g][t], offs[g + 1][t]) for g < ng, one 8-byte read per
// probe token.  Rows by grid part: the dense units run part by part, so the rows in use
// stay in L2.  A thread per posting writes offs[.][t] of the grid parts that end at it
// (those past the previous posting's part, up to its own) -- the first half of their
// entries and the second half of the entries before them -- the last posting of a list
// also those after it; a thread per term fills the entries of an empty list.  It runs
// before the planning (k_plan counts the postings of the queries that may be dense in it).
__device__ __forceinline__ void dense_off_set(uint32_t *__restrict__ offs2, uint32_t n_terms, uint32_t ng, uint32_t gg,
                                              uint32_t t, uint32_t v)
{
    if (gg < ng) offs2[2 * ((size_t)gg * n_terms + t)] = v;
    if (gg > 0) offs2[2 * ((size_t)(gg - 1) * n_terms + t) + 1] = v;
}
__global__ __launch_bounds__(256) void k_dense_offs(const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ psid,
                                                    const uint32_t *__restrict__ loff, const uint32_t *__restrict__ nE_dev,
                                                    uint32_t n_terms, uint32_t dense_w, uint32_t ng,
                                                    uint32_t *__restrict__ offs2)
{
    const uint32_t nE = *nE_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nE; i += gridDim.x * blockDim.x) {
        const uint32_t t = skeys[i], g = psid[i] / dense_w;
        const bool first = i == 0 || skeys[i - 1] != t;
        const uint32_t g0 = first ? 0u : psid[i - 1] / dense_w + 1u;   // grid parts whose first posting is i
        for (uint32_t gg = g0; gg <= g; gg++) dense_off_set(offs2, n_terms, ng, gg, t, i);
        if (i + 1 == nE || skeys[i + 1] != t)
            for (uint32_t gg = g + 1; gg <= ng; gg++) dense_off_set(offs2, n_terms, ng, gg, t, i + 1);
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_terms; t += gridDim.x * blockDim.x) {
        const uint32_t l0 = loff[t];
        if (l0 != loff[t + 1]) continue;
        for (uint32_t gg = 0; gg < ng; gg++) ((uint2 *)offs2)[(size_t)gg * n_terms + t] = make_uint2(l0, l0);
    }
}
constexpr unsigned kDenseOffsGrid = 8192;   // k_dense_offs blocks (grid-stride)

// the dense queries in ascending order: dl[doff[q]] = q where dflag[q] (doff: their scan)
__global__ void k_dense_list(const uint32_t *__restrict__ dflag, const uint32_t *__restrict__ doff, int64_t n,
                             const unsigned long long *__restrict__ ndparts, uint32_t *__restrict__ dl)
{
    if (*ndparts == 0) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && dflag[q]) dl[doff[q]] = (uint32_t)q;
}

// Dense work units, in grid-part-major order so that the wavefronts running at a time
// share one grid part's postings and signatures in L2.  dl lists the ndq dense queries in
// ascending order (a scan of k_plan's flags); their window ends are non-decreasing (every
// dense job is in size order, and a job's windows end at its own end), so the dense
// queries whose window meets grid part g are the entries [lo_g, end_g) of dl: lo_g the
// first whose window ends past g * dense_w, end_g the first at or past (g + 1) * dense_w
// - 1 (a query lies before its window).  Each part's entries are cut into tiles of
// kDenseTile queries: qlo[g] = lo_g and ntl[g] = its tiles (scanned into the tile offsets).
constexpr uint32_t kDenseTile = 16;   // queries per dense tile (one queue claim)
__global__ __launch_bounds__(256) void k_dense_tiles(const uint32_t *__restrict__ dl, const uint32_t *__restrict__ ndq_dev,
                                                     const uint32_t *__restrict__ hi, uint32_t ng, uint32_t dense_w,
                                                     const unsigned long long *__restrict__ ndparts,
                                                     uint32_t *__restrict__ qlo, uint64_t *__restrict__ ntl)
{
    const uint32_t ndq = *ndparts == 0 ? 0u : *ndq_dev;   // no dense parts: no tiles
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= ng; g += gridDim.x * blockDim.x) {
        uint32_t lo = 0, top = ndq;   // first entry whose window ends past g * dense_w
        const uint64_t start = (uint64_t)g * dense_w;
        while (lo < top) {
            const uint32_t mid = (lo + top) >> 1;
            if ((uint64_t)hi[dl[mid]] > start) top = mid;
            else lo = mid + 1;
        }
        uint32_t e = lo, etop = ndq;   // first entry at or past (g + 1) * dense_w - 1
        const uint64_t stop = (uint64_t)(g + 1u) * dense_w - 1u;
        while (e < etop) {
            const uint32_t mid = (e + etop) >> 1;
            if ((uint64_t)dl[mid] >= stop) etop = mid;
            else e = mid + 1;
        }
        qlo[g] = lo;
        ntl[g] = g < ng ? (e - lo + kDenseTile - 1u) / kDenseTile : 0u;
    }
}

// The smallest partner length ls <= l with size_ok(l, ls).  The partner lengths that
// pass the 4-way size filter form an interval containing l (each of its four bounds
// is monotone in ls), so on [0, l] the filter is false below the answer and true from
// it on: a binary search finds it.
__device__ __forceinline__ uint32_t partner_lmin(uint32_t l, float t)
{
    uint32_t lo = 0, hi = l;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (fgss_size_ok(l, mid, t)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// The largest partner length ls >= l (below 2^16) with size_ok(l, ls): the other end of
// the interval partner_lmin starts.
__device__ __forceinline__ uint32_t partner_lmax(uint32_t l, float t)
{
    uint32_t lo = l, hi = 65535u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (fgss_size_ok(l, mid, t)) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// First index in [lo, hi) whose record id is >= v, within one posting list (ids
// ascending).  Two levels: fence[f] = psid[f * kFence] for every aligned block of
// kFence postings (small, stays in L2) narrows the search to one block, which is
// one 128-byte line of psid.
__device__ __forceinline__ uint32_t lower_bound_sid(const uint32_t *__restrict__ psid, const uint32_t *__restrict__ fence,
                                                    uint32_t lo, uint32_t hi, uint32_t v)
{
    if (lo >= hi) return lo;
    // fence entries whose positions lie in [lo, hi)
    uint32_t flo = (lo + kFence - 1) / kFence, fhi = (hi - 1) / kFence + 1;
    while (flo < fhi) {                  // first fence entry with id >= v
        uint32_t mid = (flo + fhi) >> 1;
        if (fence[mid] < v) flo = mid + 1;
        else fhi = mid;
    }
    // the answer lies in [max(lo, (flo-1)*kFence), min(hi, flo*kFence)]
    uint32_t a = flo > 0 ? (flo - 1) * kFence : 0, b = flo * kFence;
    if (a < lo) a = lo;
    if (b > hi) b = hi;
    while (a < b) {
        uint32_t mid = (a + b) >> 1;
        if (psid[mid] < v) a = mid + 1;
        else b = mid;
    }
    return a;
}

// lower_bound_sid for two (range, value) pairs at once: the two searches advance
// together, so that their dependent loads overlap
__device__ __forceinline__ void lower_bound_sid2(const uint32_t *__restrict__ psid, const uint32_t *__restrict__ fence,
                                                 uint32_t lo1, uint32_t hi1, uint32_t v1, uint32_t lo2, uint32_t hi2,
                                                 uint32_t v2, uint32_t &r1, uint32_t &r2)
{
    // first fence entry with id >= v inside each range (an empty range stays empty)
    uint32_t a1 = (lo1 + kFence - 1) / kFence, b1 = lo1 < hi1 ? (hi1 - 1) / kFence + 1 : a1;
    uint32_t a2 = (lo2 + kFence - 1) / kFence, b2 = lo2 < hi2 ? (hi2 - 1) / kFence + 1 : a2;
    while (a1 < b1 || a2 < b2) {
        const uint32_t m1 = (a1 + b1) >> 1, m2 = (a2 + b2) >> 1;
        const uint32_t x1 = a1 < b1 ? fence[m1] : 0u, x2 = a2 < b2 ? fence[m2] : 0u;
        if (a1 < b1) { if (x1 < v1) a1 = m1 + 1; else b1 = m1; }
        if (a2 < b2) { if (x2 < v2) a2 = m2 + 1; else b2 = m2; }
    }
    // each answer lies in [max(lo, (a-1)*kFence), min(hi, a*kFence)]
    uint32_t p1 = max(a1 > 0 ? (a1 - 1) * kFence : 0u, lo1), q1 = min(a1 * kFence, hi1);
    uint32_t p2 = max(a2 > 0 ? (a2 - 1) * kFence : 0u, lo2), q2 = min(a2 * kFence, hi2);
    if (lo1 >= hi1) p1 = q1 = lo1;
    if (lo2 >= hi2) p2 = q2 = lo2;
    while (p1 < q1 || p2 < q2) {
        const uint32_t m1 = (p1 + q1) >> 1, m2 = (p2 + q2) >> 1;
        const uint32_t x1 = p1 < q1 ? psid[m1] : 0u, x2 = p2 < q2 ? psid[m2] : 0u;
        if (p1 < q1) { if (x1 < v1) p1 = m1 + 1; else q1 = m1; }
        if (p2 < q2) { if (x2 < v2) p2 = m2 + 1; else q2 = m2; }
    }
    r1 = p1;
    r2 = p2;
}

// First index in [lo, hi) whose record id is >= v when it is expected close to lo:
// doubling steps from lo, then a binary search inside the last step (a streamed
// item moves each cursor by a few postings per part).  It reads the postings
// themselves, whose lines the previous part has just brought in.
__device__ __forceinline__ uint32_t gallop_post(const uint2 *__restrict__ post, uint32_t lo, uint32_t hi, uint32_t v)
{
    if (lo >= hi || post[lo].x >= v) return lo;
    uint32_t prev = lo, step = 1;   // post[prev].x < v
    for (;;) {
        const uint32_t nxt = prev + step;
        if (nxt >= hi) break;
        if (post[nxt].x >= v) { hi = nxt; break; }
        prev = nxt;
        step <<= 1;
    }
    uint32_t a = prev + 1;
    while (a < hi) {
        const uint32_t mid = (a + hi) >> 1;
        if (post[mid].x < v) a = mid + 1;
        else hi = mid;
    }
    return a;
}

// first index in [lo, hi) with a[index] >= v (a ascending)
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *__restrict__ a, uint32_t lo, uint32_t hi, uint32_t v)
{
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// lanes per query in k_plan: a wavefront plans kPlanGroup queries at once (probe
// prefixes are often shorter than a wavefront)
constexpr int kPlanGroup = 16;

// Groups of kPlanGroup lanes loop over queries: the source window [q+1, hi) and, for
// each probe-prefix token, the range of its posting list inside the window (stored
// for k_join), whose total E(q) sets the number of work items.
__global__ __launch_bounds__(256) void k_plan(const uint32_t *__restrict__ tok, const uint32_t *__restrict__ pos,
                                              const uint32_t *__restrict__ len, int64_t n,
                                              const uint32_t *__restrict__ job, const float *__restrict__ tj,
                                              const uint32_t *__restrict__ jend,
                                              const uint32_t *__restrict__ probelen,
                                              const uint32_t *__restrict__ poff,
                                              const uint32_t *__restrict__ loff, const uint32_t *__restrict__ psid,
                                              const uint32_t *__restrict__ fence, uint32_t target,
                                              uint32_t heavy_target,
                                              uint2 *__restrict__ rng, uint32_t *__restrict__ hi_out,
                                              uint32_t *__restrict__ nparts, uint32_t *__restrict__ eq_out,
                                              uint64_t *__restrict__ bcnt, uint32_t stream_parts, uint32_t dense_w,
                                              uint32_t dense_pct, unsigned long long *__restrict__ ndparts,
                                              uint32_t block_records, uint32_t *__restrict__ dflag,
                                              const uint2 *__restrict__ doffs, uint32_t dterms)
{
    uint32_t dparts = 0;   // dense parts of this lane's queries
    const int64_t ngroups = (int64_t)gridDim.x * (blockDim.x / kPlanGroup);
    const uint32_t gl = (uint32_t)lane_id() & (kPlanGroup - 1);   // lane within the group
    // every group of a wavefront runs the same number of iterations (the loop bound is
    // rounded up to whole wavefronts), so the group sum below sees all its lanes
    const int64_t nround = (n + kWave / kPlanGroup - 1) / (kWave / kPlanGroup) * (kWave / kPlanGroup);
    for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kPlanGroup; g < nround; g += ngroups) {
        const bool live = g < n;
        const int64_t q = live ? g : n - 1;
        const uint32_t lq = len[q];
        const uint32_t qp = live ? probelen[q] : 0u;
        const uint32_t jb = job[q];
        // window: the s > q of the same job with len[s] >= the smallest partner length
        // when lengths are non-increasing within the job; every s > q of the job
        // otherwise (its items then apply the size filter posting by posting)
        const uint32_t je = jend[jb];
        const bool unsorted = (je & kJobUnsorted) != 0;
        const uint32_t need = partner_lmin(lq, tj[jb]);
        uint32_t lo = (uint32_t)q + 1, hi = je & ~kJobUnsorted;   // first s with len[s] < need
        if (unsorted) lo = hi;
        if (qp == 0) hi = lo;   // no probe (outside the run's queries or job group): no window
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (len[mid] >= need) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t wend = lo;   // window is [q+1, wend)
        const uint32_t ro = poff[q];
        // parts of e postings: ceil(e / target); a query of more than kHeavyParts parts
        // fills its parts fuller (fewer parts -- boundary rows, streamed rounds -- outweigh
        // the longer probe chains there); a part spans at most block_records sources
        auto parts_of = [&](uint32_t e) -> uint32_t {
            uint32_t np = e == 0 ? 0u : (e + target - 1) / target;
            if (np > kHeavyParts) np = (e + heavy_target - 1) / heavy_target;
            if (block_records != 0 && e != 0) np = max(np, (wend - (uint32_t)q - 1u + block_records - 1u) / block_records);
            return np;
        };
        // dense parts: a window of several parts whose postings fill dense parts of
        // dense_w consecutive sources to dense_pct% of their slots on average, for a probe
        // that fits the cursor slots (windows this full would need hash-table parts of a
        // few postings per probe token).  Dense parts follow the global grid of dense_w
        // records (the window's first and last parts are clipped to it), and the postings
        // counted for the test are those of the grid parts the window meets, two entries
        // of the per-part offset table per probe token: a dense query needs no searches
        // here (its units narrow the first and last part themselves).
        const uint32_t win = wend - (uint32_t)q - 1u;
        const uint32_t g0 = dense_w ? ((uint32_t)q + 1u) / dense_w : 0u, g1 = dense_w && win ? (wend - 1u) / dense_w : 0u;
        const uint32_t ndp = dense_w && win ? g1 - g0 + 1u : 0u;
        bool dense = false;
        uint32_t e = 0;
        if (!unsorted && ndp > 1 && qp < kStreamQp) {   // uniform over the group
            const uint2 *r0 = doffs + (size_t)g0 * dterms, *r1 = doffs + (size_t)g1 * dterms;
            for (uint32_t j = gl; j < qp; j += kPlanGroup) {
                const uint32_t term = tok[pos[q] + j];
                e += r1[term].y - r0[term].x;
            }
#pragma unroll
            for (int d = kPlanGroup / 2; d > 0; d >>= 1) e += __shfl_xor(e, d, kPlanGroup);
            dense = parts_of(e) > 1 && (uint64_t)e * 100u >= (uint64_t)ndp * dense_w * dense_pct;
        }
        if (!dense) {   // the posting range of each probe token inside the window
            e = 0;
            for (uint32_t j = gl; j < qp; j += kPlanGroup) {
                uint32_t term = tok[pos[q] + j];
                uint32_t L0 = loff[term], L1 = loff[term + 1];
                uint32_t a, b;   // window [q + 1, wend), wend >= q + 1
                lower_bound_sid2(psid, fence, L0, L1, (uint32_t)q + 1, L0, L1, wend, a, b);
                rng[ro + j] = make_uint2(a, b);
                e += b - a;
            }
#pragma unroll
            for (int d = kPlanGroup / 2; d > 0; d >>= 1) e += __shfl_xor(e, d, kPlanGroup);
        }
        if (live && gl == 0) {
            hi_out[q] = wend;
            eq_out[q] = e;
            const uint32_t np = dense ? 0u : parts_of(e);
            nparts[q] = dense ? (ndp | kDenseParts) : np;
            dparts += dense ? ndp : 0u;
            if (dflag) dflag[q] = dense ? 1u : 0u;   // scanned into the list of dense queries
            // internal boundaries x probe tokens, for queries whose parts read a planned
            // boundary table (heavier ones are streamed when their probe fits a cursor slot)
            bcnt[q] = !dense && np > 1 && (np <= stream_parts || qp > kStreamQp) ? (uint64_t)(np - 1) * qp : 0ull;
        }
    }
    unsigned long long w = dparts;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) w += __shfl_xor(w, d, kWave);
    if (lane_id() == 0 && w) atomicAdd(ndparts, w);
}

// sid of the start of part p of a window [a0, a0 + W) split into np parts
__device__ __forceinline__ uint32_t part_start(uint64_t a0, uint64_t W, uint32_t p, uint32_t np)
{
    return (uint32_t)(a0 + W * p / np);
}

// For every query split into np > 1 parts: the position of each internal part
// boundary in each probe token's posting range, computed once here so that k_join
// reads its ranges instead of searching.  The table is cut into chunks of
// kBoundsChunk entries, a wavefront each, whatever queries they belong to (a long
// query's thousands of entries would otherwise keep one wavefront busy while the
// others idle); each lane finds its entry's query among the chunk's and does one
// independent binary search.  Offsets are 64-bit:
// a query whose rows do not fit the table's capacity is not planned (k_items applies
// the same test and its items search their ranges instead).
// entries of the boundary table a wavefront of k_bounds takes at a time
constexpr uint32_t kBoundsChunk = 256;

// last q in [lo, hi] with ofs[q] <= g (ofs non-decreasing, ofs[lo] <= g)
__device__ __forceinline__ int64_t last_le(const uint64_t *__restrict__ ofs, int64_t lo, int64_t hi, uint64_t g)
{
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (ofs[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_bounds(int64_t n, const uint32_t *__restrict__ probelen,
                                                const uint32_t *__restrict__ poff, const uint2 *__restrict__ rng,
                                                const uint32_t *__restrict__ wlo, const uint32_t *__restrict__ hi,
                                                const uint32_t *__restrict__ nparts,
                                                const uint64_t *__restrict__ bofs, const uint64_t *__restrict__ bcnt,
                                                const uint64_t *__restrict__ nb_total,
                                                const uint32_t *__restrict__ psid,
                                                const uint32_t *__restrict__ fence, uint32_t *__restrict__ bounds,
                                                uint32_t cap_bounds)
{
    // entries at or beyond cap_bounds belong to queries that are not planned
    const uint64_t NB = min(*nb_total, (uint64_t)cap_bounds);
    const uint64_t nchunks = (NB + kBoundsChunk - 1) / kBoundsChunk;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; c < nchunks; c += nwaves) {
        const uint64_t g0 = c * kBoundsChunk, g1 = min(g0 + kBoundsChunk, NB);
        // the queries holding the chunk's first and last entries (wave-uniform)
        const int64_t qa = last_le(bofs, 0, n - 1, g0), qb = last_le(bofs, qa, n - 1, g1 - 1);
        for (uint64_t g = g0 + lane_id(); g < g1; g += kWave) {
            const int64_t q = last_le(bofs, qa, qb, g);   // its entries [bofs[q], bofs[q] + bcnt[q]) hold g
            const uint32_t np = nparts[q], qp = probelen[q];
            if (item_mode(np, qp, bcnt[q], bofs[q], cap_bounds) != IM_PLANNED) continue;
            const uint64_t a0 = wlo ? (uint64_t)wlo[q] : (uint64_t)q + 1, W = (uint64_t)hi[q] - a0;
            // token-major rows: entry k = j * (np - 1) + (p - 1), so the lanes of a
            // wavefront search one token's list for consecutive boundaries and share
            // the fence and posting lines they read
            const uint32_t k = (uint32_t)(g - bofs[q]);
            const uint32_t j = k / (np - 1), p = k - j * (np - 1) + 1;
            const uint2 r = rng[poff[q] + j];
            bounds[g] = lower_bound_sid(psid, fence, r.x, r.y, part_start(a0, W, p, np));
        }
    }
}

// Items per query (before the scan of their offsets): one streamed item per chunk of
// parts, np planned or searched parts, one whole-window item, or none.  Queries joined in
// dense parts have no items: the dense role of the join walks their (grid part, query)
// units instead.
__global__ void k_nitems(const uint32_t *__restrict__ nparts, const uint32_t *__restrict__ probelen,
                         const uint64_t *__restrict__ bofs, const uint64_t *__restrict__ bcnt, uint32_t cap_bounds,
                         int64_t n, uint32_t *__restrict__ nitems)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        const uint32_t np = nparts[q];
        const uint32_t m = item_mode(np, probelen[q], bcnt[q], bofs[q], cap_bounds);
        nitems[q] = m == IM_NONE || m == IM_DENSE ? 0u : m == IM_WHOLE ? 1u
                  : m == IM_STREAMED ? (np + kStreamChunk - 1) / kStreamChunk : np;
    }
}

// wlo: window start per query (null: q + 1, the self-join window).  rs: the queries
// are R-S half-probes, 2r (sources longer than probe r) and 2r + 1 (the others).
// A thread per item: item o belongs to the last query q with partoff[q] <= o (the
// queries without items share the offset of the next one) and is its part
// o - partoff[q], so that a query of thousands of parts is written by thousands of
// threads.
constexpr unsigned kItemsGrid = 1024;   // k_items blocks (the item count is known on the device only)

__global__ __launch_bounds__(256) void k_items(int rs, const uint32_t *__restrict__ wlo, const uint32_t *__restrict__ hi,
                                               const uint32_t *__restrict__ nparts, const uint32_t *__restrict__ probelen,
                                               const uint32_t *__restrict__ eq, const uint32_t *__restrict__ partoff,
                                               const uint64_t *__restrict__ bofs, const uint64_t *__restrict__ bcnt,
                                               uint32_t cap_bounds, const uint32_t *__restrict__ job,
                                               const float *__restrict__ tj, const uint32_t *__restrict__ len,
                                               const uint32_t *__restrict__ pos, const uint32_t *__restrict__ poff,
                                               int64_t n, const uint32_t *__restrict__ n_items_dev,
                                               Item *__restrict__ items, uint32_t cap_items,
                                               const uint32_t *__restrict__ jend)
{
    const uint32_t total = *n_items_dev;
    if (total > cap_items) return;   // too small: the host resizes and reruns
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < total; o += gridDim.x * blockDim.x) {
        int64_t lo = 0, top = n - 1;   // partoff[0] = 0 <= o
        while (lo < top) {
            const int64_t mid = (lo + top + 1) >> 1;
            if (partoff[mid] <= o) lo = mid;
            else top = mid - 1;
        }
        const int64_t q = lo;
        const uint32_t p = o - partoff[q];
        const uint32_t np = nparts[q];
        const uint32_t mode = item_mode(np, probelen[q], bcnt[q], bofs[q], cap_bounds);
        const uint64_t a0 = wlo ? (uint64_t)wlo[q] : (uint64_t)q + 1, W = (uint64_t)hi[q] - a0;
        const uint32_t per = min((eq[q] + np - 1) / np, kPadMask);   // expected postings per part: sizes its LDS table
        const int64_t r = rs ? q >> 1 : q;           // the record whose tokens probe
        const uint32_t flag = rs && (q & 1) == 0 ? kMidProbe : 0u;
        const uint32_t fl = mode == IM_WHOLE ? kWholeWindow : mode == IM_PLANNED ? kPlanned
                          : mode == IM_STREAMED ? kStreamed : 0u;
        Item it;
        it.q = (uint32_t)r;
        if (mode == IM_STREAMED) {   // chunk p: parts [p * kStreamChunk, ...) of the query's np
            const uint32_t p0 = p * kStreamChunk, p1 = min(p0 + kStreamChunk, np);
            it.a = part_start(a0, W, p0, np);
            it.b = part_start(a0, W, p1, np);
            it.nparts = p1 - p0;
        } else {
            it.a = part_start(a0, W, p, np);
            it.b = part_start(a0, W, p + 1, np);
            it.nparts = np;
        }
        it.pad = per | fl;
        it.part = p;
        it.bofs = mode == IM_PLANNED ? (uint32_t)bofs[q] : 0u;
        it.job = job[r];
        it.t = tj[it.job];
        it.lq = len[r] | flag | (jend != nullptr && (jend[it.job] & kJobUnsorted) != 0 ? kSizeFilter : 0u);
        it.qpos = pos[r];
        it.qpoff = poff[q];
        items[o] = it;
    }
}

// split every overflowed item -- or, for a streamed item, its overflowed part
// (ovfp) -- into `ways` equal source sub-windows whose parts search their ranges
__global__ void k_resplit(const Item *__restrict__ in, const uint32_t *__restrict__ ovf,
                          const uint32_t *__restrict__ ovfp, uint32_t novf, uint32_t ways, Item *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= novf) return;
    Item it = in[ovf[i]];
    if (ovfp[i] != kEmpty) {
        const uint64_t W0 = (uint64_t)it.b - it.a;
        const uint32_t a = part_start(it.a, W0, ovfp[i], it.nparts), b = part_start(it.a, W0, ovfp[i] + 1, it.nparts);
        it.a = a;
        it.b = b;
    }
    uint64_t W = (uint64_t)it.b - it.a;
    for (uint32_t p = 0; p < ways; p++) {
        Item o = it;
        o.a = (uint32_t)(it.a + W * p / ways);
        o.b = (uint32_t)(it.a + W * (p + 1) / ways);
        o.pad = it.pad & kPadMask;   // same table size as the parent: four times the room per posting
        o.part = 0;
        o.nparts = 0;
        o.bofs = 0;
        out[i * ways + p] = o;
    }
}

// ---------------------------------------------------------------------------
// the join kernel: one wavefront per work item, items taken from a queue
// ---------------------------------------------------------------------------
struct JoinArgs {
    const uint32_t *tok, *pos, *len;    // source records (and the queries of a self-join)
    const uint32_t *qtok;               // query tokens (R-S: the probe records)
    const uint32_t *qsig, *qsig64;      // query signatures
    const uint8_t *qsigbad;
    const uint32_t *loff;
    const uint2 *post;          // postings (record id, token position | length << 16), grouped by term
    const uint32_t *pden;       // the same postings as record | position << 24, for dense parts (null: none)
    const uint2 *doffs;         // dense parts: per grid part g < dng and term t, the first postings of t at
                                // or after records g and g + 1 x the grid width (row g, dterms entries; k_dense_offs)
    uint32_t dng;               // grid parts of the dense offset table (its rows)
    uint32_t dterms;            // terms (the offset table's row length)
    const uint32_t *dl;         // the dense queries in ascending order (k_dense_list)
    const uint32_t *ndq_dev;    // their number (device)
    const uint32_t *dqlo;       // per grid part: the first entry of dl in its dense tiles (k_dense_tiles)
    const uint64_t *dtpre;      // per grid part: dense tiles of the parts before it
    const unsigned long long *n_tiles_dev;   // dense tiles of kDenseTile units (device; null: none)
    const uint32_t *hi;         // per query: the end of its window
    const uint32_t *nparts;     // per query: its parts (kDenseParts: joined in dense units)
    int64_t n;                  // records (the queries of a self-join)
    const uint32_t *psid;       // record ids of the postings alone
    const uint32_t *fence;      // psid[f * kFence]
    const uint32_t *sig;        // per record: kSigWords words of byte counts per token bucket
    const uint32_t *sig64;      // per record: kSig64Words words, 64 merged buckets
    const uint8_t *sigbad;      // 1 when a bucket count did not fit in a byte
    const uint2 *rng;           // per probe token: posting range inside the query's window
    const uint32_t *poff;       // per query: first entry of its probe tokens in rng
    const uint32_t *bounds;     // planned split boundaries (see k_bounds)
    const Item *items;
    const uint32_t *n_items_dev;  // number of items (device); the launch is void if above n_items_cap
    uint32_t n_items_cap;
    uint32_t dense_blocks;      // blocks of the launch that join the dense units (the first ones)
    const uint32_t *job;        // per record: job id
    const float *tj;            // per job: Jaccard threshold
    unsigned long long *ctr_job;  // per job: kJobCtr counters
    int32_t *out;               // (q, s, overlap) triples
    uint32_t out_cap;
    unsigned long long *wctr;   // kWaveSlot words per wavefront of the grid
    uint32_t *ovf;              // indices of overflowed items
    uint32_t *ovfp;             // ... and, for streamed items, the overflowed part (else kEmpty)
    uint32_t ovf_cap;           // records beyond it are counted only: the host grows the list and reruns
    unsigned long long *ctr;
};

__device__ __forceinline__ uint32_t slot_hash(uint32_t s, int log2cap)
{
    return (s * 0x9E3779B1u) >> (32 - log2cap);
}

// the probe step of key s (odd, so that a key's probes visit every slot of a power-of-two
// table): double hashing keeps the probe chains of an item's keys apart, where linear
// probing grows clusters whose longest chain sets the rounds of a wavefront's insert
__device__ __forceinline__ uint32_t slot_step(uint32_t s, int log2cap)
{
    return ((s * 0x85EBCA6Bu) >> (32 - log2cap)) | 1u;
}

// Order LDS traffic between lanes of one wavefront: a compiler barrier plus a
// wavefront-scope fence (the hardware executes a wave's LDS operations in order).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int ceil_log2(uint32_t x)
{
    return x <= 1 ? 0 : 32 - __clz(x - 1);
}

// Phase A's table update for the kUnroll postings a lane holds (pe[k], query position
// jpos[k]; bit k of valm set when posting k takes part).  The probes of all kUnroll
// postings are issued together and wait once; collisions are resolved for all of them
// in one loop (open addressing with double hashing stays exact whatever the order in
// which keys are placed, as no slot is emptied during the phase, and every pending probe
// moves to a new slot of its key's cycle each round, so one wave-uniform round counter
// bounds them by tsize slots).  New keys are
// appended to clist.  Lanes without work aim their LDS updates at scr[lane].  Returns
// true on any lane when some posting found no free slot.
template <int N>
__device__ __forceinline__ bool table_insert(const uint2 (&pe)[N], uint32_t valm, const uint32_t (&jpos)[N],
                                             uint32_t *tkey, uint32_t *tcnt, uint32_t *tlast, uint32_t *clist,
                                             uint32_t *scr, int lane, int tlog, uint32_t tsize, uint32_t &nlist)
{
    uint32_t newm = 0, pendm = 0, failm = 0;
    uint32_t h[N], prev[N], hs[N];
#pragma unroll
    for (int k = 0; k < N; k++) h[k] = slot_hash(pe[k].x, tlog);
#pragma unroll
    for (int k = 0; k < N; k++) prev[k] = atomicCAS(((valm >> k) & 1u) ? &tkey[h[k]] : &scr[lane], kEmpty, pe[k].x);
#pragma unroll
    for (int k = 0; k < N; k++) {
        const uint32_t vk = (valm >> k) & 1u;
        newm |= (vk & (uint32_t)(prev[k] == kEmpty)) << k;
        pendm |= (vk & (uint32_t)(prev[k] != kEmpty) & (uint32_t)(prev[k] != pe[k].x)) << k;
    }
    if (__ballot(pendm != 0)) {
#pragma unroll
        for (int k = 0; k < N; k++) hs[k] = slot_step(pe[k].x, tlog);
    }
    for (uint32_t round = 1; __ballot(pendm != 0); round++) {
        if (round >= tsize) { failm |= pendm; break; }
#pragma unroll
        for (int k = 0; k < N; k++) {
            const uint32_t pk = (pendm >> k) & 1u;
            h[k] = (h[k] + (pk ? hs[k] : 0u)) & (tsize - 1u);
            prev[k] = atomicCAS(pk ? &tkey[h[k]] : &scr[lane], kEmpty, pe[k].x);
        }
#pragma unroll
        for (int k = 0; k < N; k++) {
            const uint32_t pk = (pendm >> k) & 1u;
            const uint32_t placed = pk & (uint32_t)(prev[k] == kEmpty);
            newm |= placed << k;
            pendm &= ~((placed | (pk & (uint32_t)(prev[k] == pe[k].x))) << k);
        }
    }
    const uint32_t hitm = valm & ~failm;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const bool hit = (hitm >> k) & 1u;
        atomicAdd(hit ? &tcnt[h[k]] : &scr[lane], 1u);
        atomicMax(hit ? &tlast[h[k]] : &scr[lane], (jpos[k] << 16) | (pe[k].y & 0xFFFFu));
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        const bool isnew = (newm >> k) & 1u;
        const uint64_t nm = __ballot(isnew);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(nm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
        *(isnew ? &clist[nlist + below] : &scr[lane]) = h[k] | (pe[k].y & 0xFFFF0000u);
        nlist += (uint32_t)__popcll(nm);
    }
    return __ballot(failm != 0) != 0;
}

// Work item = (query q, source window [a, b)).  Phases, all inside one wavefront:
//  A  candidate generation: the postings of the probe-prefix tokens whose record id
//     lies in the window, flattened over all tokens, kUnroll per lane in flight; per
//     source an LDS hash slot keeps the match count and the last (query position,
//     source position); each new source is appended to a compact list together with
//     its length (carried by the posting).
//  B  positional filter over the list (counted as the reference's "# Candidates"),
//     then the bound m + min(|Q|-j-1, |S|-s-1) >= floor(bound).
//  C  signature bound: overlap <= (|Q| + |S| - SAD(countsQ, countsS)) / 2 over 128
//     byte-wide token-bucket counts, one lane per candidate.
//  D  exact overlap for the rest: the source tokens after the last prefix match,
//     64 per step, each looked up in the query (staged in LDS when it fits),
//     ballot + popcount.
// Each phase compacts the list in place and resets the table slots it drops, so the
// table is empty again when the item ends: no per-item clear.
// postings each lane loads per step of phase A, and LDS table slots per expected
// posting (tuning knobs, overridable at build time)
#ifndef FGSS_UNROLL
#define FGSS_UNROLL 4
#endif
#ifndef FGSS_PHASE_CLOCKS
#define FGSS_PHASE_CLOCKS 0   // 1: the per-phase shader clocks (phase_cycles; they cost 1-4% of the join)
#endif
#ifndef FGSS_TABLE_FACTOR
#define FGSS_TABLE_FACTOR 2
#endif
constexpr int kUnroll = FGSS_UNROLL;
constexpr int kQStage = 160;   // query tokens staged in LDS for verification (longer queries stay in HBM)
// per-wave counters (lane k of one 64-bit register holds counter k): totals,
// shader-clock phases, the current job's counters
enum WaveAcc { W_ESCAN = 0, W_CAND, W_SIG, W_VER, W_SIG2, W_INTER, W_VTOK, W_RETRY, W_DENSE, W_DPOST, W_CYC0,
               W_JOB0 = W_CYC0 + 6, kAccSlots = W_JOB0 + kJobCtr };
// per-wavefront slot in global memory at the end of a launch: the kAccSlots
// counters and the wavefront's last job
constexpr int kWaveSlot = 32;
static_assert(kAccSlots + 1 <= kWaveSlot && kAccSlots + 1 <= kWave, "slot layout");
// LDS words per wavefront: table (key, count, last match), list, token ranges, query stage
constexpr uint32_t join_wave_words(int cap)
{
    return 4u * cap + 2u * kTokBatch + 1u + kQStage;
}

// RS: an R-S join against an index holding max_prefix(|S|, t_min) tokens per source;
// postings past max_prefix(|S|, t) of the item's threshold are skipped.  rblock /
// rblocks: the block's index among the launch's blocks and their number; sblock: its
// counter slot (after the dense kernel's blocks).
template <int CAP, bool RS>
__device__ __forceinline__ void join_items(const JoinArgs &A, uint32_t rblock, uint32_t rblocks, uint32_t sblock)
{
    constexpr int LOG2CAP = CAP == 256 ? 8 : CAP == 512 ? 9 : CAP == 1024 ? 10 : CAP == 2048 ? 11 : 12;
    static_assert((1 << LOG2CAP) == CAP, "CAP must be a power of two");
    extern __shared__ uint32_t smem[];
    const int w = (int)threadIdx.x / kWave;
    const int lane = lane_id();
    uint32_t *tkey = smem + (size_t)w * join_wave_words(CAP);
    uint32_t *tcnt = tkey + CAP;
    uint32_t *tlast = tcnt + CAP;
    uint32_t *clist = tlast + CAP;        // slot | |S| << 16, one per distinct source
    uint32_t *rlo = clist + CAP;
    uint32_t *rpre = rlo + kTokBatch;     // kTokBatch + 1 entries
    uint32_t *qstage = rpre + kTokBatch + 1;
    // a launch whose item list or boundary table did not fit does nothing: the host
    // sees the totals, grows the buffers and runs the pipeline again
    if (*A.n_items_dev > A.n_items_cap) return;
    const uint64_t n_items = (uint64_t)*A.n_items_dev;
    {   // a wavefront with no static chunk and no queue to claim from writes a zero
        // counter slot and leaves at once
        const uint64_t nwg = (uint64_t)rblocks * (blockDim.x / kWave);
        const uint64_t g = min((n_items + nwg - 1) / nwg, (uint64_t)kQueueBatch);
        const uint64_t gwi = (uint64_t)rblock * (blockDim.x / kWave) + (uint32_t)w;
        if (gwi * g >= n_items && nwg * g >= n_items) {
            unsigned long long *slot = A.wctr + ((size_t)sblock * join_waves(CAP) + (size_t)w) * kWaveSlot;
            if (lane < kAccSlots) slot[lane] = 0ull;
            if (lane == kAccSlots) slot[lane] = kEmpty;
            return;
        }
    }

    // the hash table starts empty
    for (uint32_t i = lane; i < CAP; i += kWave) {
        tkey[i] = kEmpty;
        tcnt[i] = 0;
        tlast[i] = 0;
    }
    wave_sync();
// shader-clock time per phase, summed per wavefront (FGSS_PHASE_CLOCKS=0 builds
// without the clock reads and their LDS updates)
#if FGSS_PHASE_CLOCKS
#define PHASE_MARK(k)                                  \
    do {                                               \
        c1 = clock64();                                \
        WACC(W_CYC0 + (k), c1 - c0);                   \
        c0 = c1;                                       \
    } while (0)
#define PHASE_START(r) const unsigned long long r = clock64()
#define PHASE_ADD(k, r) WACC(W_CYC0 + (k), clock64() - (r))
#else
#define PHASE_MARK(k) do { } while (0)
#define PHASE_START(r) do { } while (0)
#define PHASE_ADD(k, r) do { } while (0)
#endif
    // The wavefront's counters live in one 64-bit register, counter k in lane k: an
    // update (of a wave-uniform value) is a compare and a select, no branch or LDS
    // round trip.
    unsigned long long wv = 0;
// (v is evaluated on every lane before the select: a ballot inside it must see them all)
#define WACC(k, v)                                                        \
    do {                                                                  \
        const unsigned long long wacc_v_ = (unsigned long long)(v);       \
        wv += (uint32_t)lane == (uint32_t)(k) ? wacc_v_ : 0ull;           \
    } while (0)

    unsigned long long c0 = clock64(), c1 = 0;
    (void)c0;
    (void)c1;
    // Each wavefront's first g0 items are its own (consecutive, by its index in the
    // grid), g0 = ceil(items / wavefronts) up to kQueueBatch, so that a launch of few
    // items (a small served batch) spreads over every wavefront without atomics; the
    // items past the static chunks are claimed kQueueBatch at a time, one global
    // atomic per claim (claims of fewer items contend on the counter).
    const uint64_t nwaves_grid = (uint64_t)rblocks * (blockDim.x / kWave);
    const uint64_t g0 = min((n_items + nwaves_grid - 1) / nwaves_grid, (uint64_t)kQueueBatch);
    const uint64_t qbase = nwaves_grid * g0;   // first item of the dynamic claims
    const uint64_t gw = (uint64_t)rblock * (blockDim.x / kWave) + (uint32_t)w;
    uint64_t batch_next = min(gw * g0, n_items), batch_end = min(gw * g0 + g0, n_items);
    uint32_t cur_job = kEmpty;
    float cur_t = 0.0f;   // the threshold of cur_job
    // the next item of the claimed batch is loaded while the current one runs
    uint32_t nxtw = 0;
    bool have_nxt = false;
    for (;;) {
        if (batch_next >= batch_end) {
            if (qbase >= n_items) break;   // every item was in a static chunk
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(&A.ctr[C_QUEUE], (unsigned long long)kQueueBatch);
            b0 = qbase + (((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(b0 >> 32)) << 32) |
                          (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)b0));
            batch_next = min((uint64_t)b0, n_items);
            batch_end = min((uint64_t)b0 + kQueueBatch, n_items);
            have_nxt = false;
        }
        if (batch_next >= n_items) break;
        const uint64_t it = batch_next++;
        Item item;
        // an item is kItemWords words, lane k holding word k (one register for the
        // prefetched item instead of kItemWords)
        const uint32_t ldw = have_nxt ? nxtw : item_word(A.items, (uint32_t)it, lane);
        have_nxt = batch_next < batch_end;
        if (have_nxt) nxtw = item_word(A.items, (uint32_t)batch_next, lane);
        // the item is the same for every lane: its fields live in scalar registers
        item.q = __builtin_amdgcn_readlane(ldw, 0);
        item.a = __builtin_amdgcn_readlane(ldw, 1);
        item.b = __builtin_amdgcn_readlane(ldw, 2);
        item.pad = __builtin_amdgcn_readlane(ldw, 3);
        item.part = __builtin_amdgcn_readlane(ldw, 4);
        item.nparts = __builtin_amdgcn_readlane(ldw, 5);
        item.bofs = __builtin_amdgcn_readlane(ldw, 6);
        item.job = __builtin_amdgcn_readlane(ldw, 7);
        item.t = __int_as_float(__builtin_amdgcn_readlane(ldw, 8));
        item.lq = __builtin_amdgcn_readlane(ldw, 9);
        item.qpos = __builtin_amdgcn_readlane(ldw, 10);
        item.qpoff = __builtin_amdgcn_readlane(ldw, 11);
        const uint32_t q = item.q;
        const uint32_t jq = item.job;
        const float t = item.t;
        if (jq != cur_job) {   // per-job counters: flush when the job changes
            if (cur_job != kEmpty) {
#pragma unroll
                for (int k = 0; k < kJobCtr; k++) {
                    const unsigned long long v = (unsigned long long)__builtin_amdgcn_readlane((uint32_t)wv, W_JOB0 + k) |
                                                 ((unsigned long long)__builtin_amdgcn_readlane((uint32_t)(wv >> 32), W_JOB0 + k) << 32);
                    if (lane == 0 && v) atomicAdd(&A.ctr_job[(size_t)cur_job * kJobCtr + k], v);
                }
            }
            wv = lane >= W_JOB0 && lane < W_JOB0 + kJobCtr ? 0ull : wv;
            cur_job = jq;
            cur_t = t;
        }
        const uint32_t lq = item.lq & 0xFFFFu;
        const bool midprobe = RS && (item.lq & kMidProbe) != 0;
        // a job in any record order: the partner lengths that pass the 4-way size filter,
        // an interval [szlo, szhi] (each of the filter's bounds is monotone in |S|)
        const bool sizefilt = !RS && (item.lq & kSizeFilter) != 0;
        uint32_t szlo = 0, szhi = 0xFFFFu;
        if (sizefilt) {
            szlo = partner_lmin(lq, item.t);
            szhi = partner_lmax(lq, item.t);
        }
        const uint32_t qpos = item.qpos;
        uint32_t qp = midprobe ? fgss_mid_prefix(lq, t) : fgss_max_prefix(lq, t);
        if (qp > lq) qp = lq;
        // table sized for FGSS_TABLE_FACTOR x the expected postings (64 .. CAP slots)
        int tlog = ceil_log2(FGSS_TABLE_FACTOR * (item.pad & kPadMask));
        tlog = tlog < 6 ? 6 : (tlog > LOG2CAP ? LOG2CAP : tlog);
        const uint32_t tsize = 1u << tlog;

        const bool whole = (item.pad & kWholeWindow) != 0;
        const bool planned = (item.pad & kPlanned) != 0;
        const bool streamed = (item.pad & kStreamed) != 0;
        const uint2 *qrng = A.rng + item.qpoff;
        // planned: token j's boundaries are bounds[bofs + j * (nparts - 1) + p - 1], p = 1 .. nparts - 1
        const uint32_t *brow = planned ? A.bounds + item.bofs : nullptr;
        const uint32_t bstride = planned ? item.nparts - 1 : 0u;
        const bool has_lo = planned && item.part > 0, has_hi = planned && item.part + 1 < item.nparts;
        // a streamed item walks its nparts windows in order with one cursor per probe
        // token (LDS), starting at the token's first posting inside the item's window
        uint32_t *cur = rlo;
        static_assert(2 * kTokBatch + 1 >= (int)kStreamQp, "cursors live in rlo and rpre");
        if (streamed) {
            for (uint32_t j = lane; j < qp; j += kWave) {
                const uint2 r = qrng[j];
                cur[j] = item.part == 0 ? r.x : lower_bound_sid(A.psid, A.fence, r.x, r.y, item.a);
            }
            wave_sync();
        }
        const uint32_t nwin = streamed ? item.nparts : 1u;
        PHASE_MARK(0);

        // a streamed item joins its parts a window of `span` consecutive parts at a
        // time: the span doubles while the windows leave the table at most 40% full
        // and halves when one fills it more than 70% (fewer, fuller windows: fewer
        // rounds and signature passes, posting lines read once); a window of several
        // parts that overflows runs again part by part
        uint32_t span = 1;
        for (uint32_t p0 = 0, p1 = 0; p0 < nwin; p0 = p1) {
        p1 = min(p0 + span, nwin);
        uint32_t wa = item.a, wb = item.b;
        if (streamed) {
            const uint64_t W = (uint64_t)item.b - item.a;
            wa = part_start(item.a, W, p0, nwin);
            wb = part_start(item.a, W, p1, nwin);
        }

        // ---- A: candidate generation (filtering.cu:94-139) ----
        bool overflow = false;
        uint32_t escan = 0, nlist = 0, jdone = 0;
        for (uint32_t jb = 0; !streamed && jb < qp && !overflow; jb += kTokBatch) {
            const uint32_t nb = min((uint32_t)kTokBatch, qp - jb);
            PHASE_START(r0);
            for (uint32_t jj = lane; jj < nb; jj += kWave) {
                const uint2 r = qrng[jb + jj];   // the token's postings inside the whole window
                uint32_t lo = r.x, hi = r.y;
                if (planned) {                   // boundaries precomputed by k_bounds
                    const uint32_t *bj = brow + (size_t)(jb + jj) * bstride + item.part;
                    if (has_lo) lo = bj[-1];
                    if (has_hi) hi = bj[0];
                } else if (!whole) {             // a re-split item: narrow by search
                    lo = lower_bound_sid(A.psid, A.fence, r.x, r.y, wa);
                    hi = lower_bound_sid(A.psid, A.fence, lo, r.y, wb);
                }
                rlo[jj] = lo;
                rpre[jj] = hi - lo;
            }
            jdone = jb + nb;
            wave_sync();
            // exclusive scan of the segment lengths, kTokBatch/64 segments per lane
            // (segments lane * PER + k), kept in registers: slen their lengths, sbeg
            // their flattened starts; rlo[jj] becomes the posting index of flattened
            // position 0 in segment jj's numbering, so that posting e of segment jj is
            // post[rlo[jj] + e]
            constexpr int PER = kTokBatch / kWave;
            uint32_t slen[PER], sbeg[PER], sum = 0;
#pragma unroll
            for (int k = 0; k < PER; k++) {
                uint32_t jj = (uint32_t)lane * PER + k;
                slen[k] = jj < nb ? rpre[jj] : 0u;
                sum += slen[k];
            }
            uint32_t inc = wave_inclusive_scan_dpp(sum);
            uint32_t run = inc - sum;
#pragma unroll
            for (int k = 0; k < PER; k++) {
                uint32_t jj = (uint32_t)lane * PER + k;
                sbeg[k] = slen[k] != 0 ? run : ~0u;   // an empty segment marks nothing
                if (jj < nb) rlo[jj] -= run;
                run += slen[k];
            }
            // wave-uniform: the posting loop below is a scalar loop
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
            wave_sync();
            PHASE_ADD(5, r0);
            escan += total;
            // The posting loop runs without branches on per-lane conditions (on this
            // chip every divergent branch costs scalar issue slots, which bound the
            // kernel): lanes without a posting work on the last one and aim their LDS
            // updates at a word of their own in the idle query stage.
            uint32_t *scr = qstage;
            // segment marks of a step, one byte per flattened position, after the
            // per-lane words in the query stage
            uint8_t *smark = (uint8_t *)(qstage + kWave);
            static_assert(kUnroll * kWave + 1 <= 4 * (kQStage - kWave), "a step's marks and a spare byte fit the query stage");
            uint32_t segc = 0;   // segment of the last posting of the previous step
            // A step takes N postings per lane (rows of kWave): full steps of kUnroll
            // rows, then a tail of 2 or 1 rows, so that a short range (most items hold
            // a few hundred postings or less) does not pay for empty rows.
            auto fstep = [&](auto rows, uint32_t e0) -> bool {
                constexpr int N = decltype(rows)::value;
                uint2 pe[N];
                uint32_t seg[N], e[N];
                // lanes past the end work on the last posting (a real one: every load
                // and LDS read below is unconditional) and take no part in the updates
#pragma unroll
                for (int k = 0; k < N; k++) e[k] = min(e0 + (uint32_t)(k * kWave) + lane, total - 1u);
                // segment of each flattened posting: the last non-empty segment starting
                // at or before it.  Non-empty segments start at distinct positions; each
                // one starting inside this step marks its position with its index, and a
                // max-scan over the step's rows (carrying segc in from the step before)
                // hands every posting its segment.
                static_assert(kTokBatch <= 256, "segment indices fit a byte");
                ((uint32_t *)smark)[lane] = 0u;
#pragma unroll
                for (int k = 0; k < PER; k++)   // starts outside the step, and empty segments, mark a spare byte
                    smark[min(sbeg[k] - e0, (uint32_t)(N * kWave))] = (uint8_t)(lane * PER + k);
                wave_sync();
                uint32_t c = segc;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const uint32_t m = wave_max_scan((uint32_t)smark[k * kWave + lane]);
                    seg[k] = max(m, c);
                    c = max(c, (uint32_t)__builtin_amdgcn_readlane((int)m, kWave - 1));
                }
                segc = c;
#pragma unroll
                for (int k = 0; k < N; k++) pe[k] = A.post[rlo[seg[k]] + e[k]];
                // bit k: posting k is in range and, for R-S items, inside the source's prefix
                uint32_t valm = 0, jpos[N];
#pragma unroll
                for (int k = 0; k < N; k++) {
                    uint32_t v = (e0 + (uint32_t)(k * kWave) + (uint32_t)lane) < total;
                    if (RS) {   // the source's own prefix at t
                        const uint32_t ls = pe[k].y >> 16;
                        v &= (uint32_t)((pe[k].y & 0xFFFFu) < (midprobe ? fgss_max_prefix(ls, t) : fgss_mid_prefix(ls, t)));
                    }
                    if (sizefilt) {   // postings of non-partners are neither scanned nor counted
                        const uint32_t ok = (uint32_t)((pe[k].y >> 16) - szlo <= szhi - szlo);
                        escan -= (uint32_t)__popcll(__ballot(v & (ok ^ 1u)));
                        v &= ok;
                    }
                    valm |= v << k;
                    jpos[k] = jb + seg[k];
                }
                return table_insert(pe, valm, jpos, tkey, tcnt, tlast, clist, scr, lane, tlog, tsize, nlist);
            };
            for (uint32_t e0 = 0; e0 < total;) {
                const uint32_t rem = total - e0;   // wave-uniform
                bool full;
                if (rem > 3u * kWave) { full = fstep(std::integral_constant<int, kUnroll>(), e0); e0 += kWave * kUnroll; }
                else if (rem > (uint32_t)kWave) { full = fstep(std::integral_constant<int, 2>(), e0); e0 += 2 * kWave; }
                else { full = fstep(std::integral_constant<int, 1>(), e0); e0 += kWave; }
                if (full) {
                    overflow = true;
                    break;
                }
            }
            wave_sync();
        }
        if (streamed) {
            // A streamed part [wa, wb): every probe token with postings left takes the
            // next K of them from its cursor (independent loads, no search).  Those
            // before wb are this part's; the group of K lanes of a token moves its
            // cursor past them, and a token whose K were all before wb takes K more
            // (K doubling up to 64) in the next round.
            uint8_t *toks = (uint8_t *)qstage;          // tokens of a round (bytes)
            uint32_t *satm = qstage + kStreamQp / 4;    // kStreamQp bits: tokens going on to the next round
            uint32_t *sscr = satm + kStreamQp / 32;     // a word per lane for updates without work
            static_assert(kStreamQp / 4 + kStreamQp / 32 + kWave <= kQStage, "token list fits the query stage");
            // one pass over the tokens toks[0, total / K), K = 1 << logK lanes each
            uint32_t logK = 3, K = 8, total = 0;
            // rows per step as in the flattened loop
            auto sstep = [&](auto rows, uint32_t e0) -> bool {
                constexpr int N = decltype(rows)::value;
                // every lane reads a posting (lanes past the end the last group's;
                // reads past a token's list land in the next list or the spare
                // postings) together with its token's list end, and then tests it
                uint2 pe[N];
                uint32_t jv[N], curj[N], idx[N], end[N];
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const uint32_t ec = min(e0 + (uint32_t)(k * kWave) + lane, total - 1u);
                    jv[k] = toks[ec >> logK];
                    curj[k] = cur[jv[k]];
                    idx[k] = curj[k] + (ec & (K - 1));
                    pe[k] = A.post[idx[k]];
                    end[k] = qrng[jv[k]].y;
                }
                uint32_t valm = 0;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const uint32_t e = e0 + (uint32_t)(k * kWave) + lane;
                    const bool below = e < total && idx[k] < end[k] && pe[k].x < wb;
                    // below is a prefix of each token's K lanes (postings ascend by
                    // record): its length is the count, taken by the group's first lane
                    const uint64_t bm = __ballot(below);
                    escan += (uint32_t)__popcll(bm);
                    const uint32_t gs = (uint32_t)lane & ~(K - 1);
                    const uint64_t gmask = K >= 64 ? ~0ull : ((1ull << K) - 1ull);
                    const uint32_t cnt = (uint32_t)__popcll((bm >> gs) & gmask);
                    // each token has one group per pass, in one row k: the groups of a
                    // row move different cursors
                    const bool lead = e < total && ((uint32_t)lane & (K - 1)) == 0 && cnt > 0;
                    *(lead ? &cur[jv[k]] : &sscr[lane]) = curj[k] + cnt;
                    atomicOr(lead && cnt == K ? &satm[jv[k] >> 5] : &sscr[lane], 1u << (jv[k] & 31));
                    uint32_t v = below;
                    if (RS) {   // the source's own prefix at t
                        const uint32_t ls = pe[k].y >> 16;
                        v &= (uint32_t)((pe[k].y & 0xFFFFu) < (midprobe ? fgss_max_prefix(ls, t) : fgss_mid_prefix(ls, t)));
                    }
                    if (sizefilt) {   // the cursor still moves past non-partners (bm above)
                        const uint32_t ok = (uint32_t)((pe[k].y >> 16) - szlo <= szhi - szlo);
                        escan -= (uint32_t)__popcll(__ballot(v & (ok ^ 1u)));
                        v &= ok;
                    }
                    valm |= v << k;
                }
                return table_insert(pe, valm, jv, tkey, tcnt, tlast, clist, sscr, lane, tlog, tsize, nlist);
            };
            auto pass = [&]() -> bool {   // true when the hash table filled up
                constexpr int NS = kUnroll;
                for (uint32_t e0 = 0; e0 < total;) {
                    const uint32_t rem = total - e0;   // wave-uniform
                    bool full;
                    if (rem > (uint32_t)(NS - 1) * kWave) { full = sstep(std::integral_constant<int, NS>(), e0); e0 += kWave * NS; }
                    else if (rem > (uint32_t)kWave) { full = sstep(std::integral_constant<int, 2>(), e0); e0 += 2 * kWave; }
                    else { full = sstep(std::integral_constant<int, 1>(), e0); e0 += kWave; }
                    if (full) return true;
                }
                return false;
            };
            {
                uint32_t nact = 0;
                for (uint32_t j0 = 0; j0 < qp; j0 += kWave) {
                    const uint32_t j = j0 + lane;
                    const bool act = j < qp && cur[j] < qrng[j].y;
                    const uint64_t m = __ballot(act);
                    if (act) toks[nact + (uint32_t)__popcll(m & lanemask_lt())] = (uint8_t)j;
                    nact += (uint32_t)__popcll(m);
                }
                wave_sync();
                // first round: K = 8 lanes per token.  A round waits for the cursor moves of
                // the one before it, so fewer rounds beat fewer idle lanes up to a point
                // (1M long records at t = 0.8, about 2.4 postings per token and part:
                // K = 4 2.11 s, 8 2.02 s, 16 2.51 s; starting at 1 2.59 s).
                logK = 3;
                while (nact > 0 && !overflow) {
                    if (lane < (int)(kStreamQp / 32)) satm[lane] = 0;
                    wave_sync();
                    K = 1u << logK;
                    total = nact << logK;
                    overflow = pass();
                    wave_sync();
                    nact = 0;
                    for (uint32_t j0 = 0; j0 < qp; j0 += kWave) {
                        const uint32_t j = j0 + lane;
                        const bool act = j < qp && ((satm[j >> 5] >> (j & 31)) & 1u) != 0;
                        const uint64_t m = __ballot(act);
                        wave_sync();
                        if (act) toks[nact + (uint32_t)__popcll(m & lanemask_lt())] = (uint8_t)j;
                        nact += (uint32_t)__popcll(m);
                    }
                    wave_sync();
                    if (logK < 6) logK++;
                }
            }
        }
        if (overflow) {
            for (uint32_t k = lane; k < nlist; k += kWave) {
                const uint32_t h = clist[k] & 0xFFFFu;
                tkey[h] = kEmpty; tcnt[h] = 0; tlast[h] = 0;
            }
            wave_sync();
            span = 1;
            if (streamed && p1 - p0 > 1) {
                // a window of several parts runs again part by part: cursors back to
                // the first posting of each token at or after the window's start
                WACC(W_RETRY, 1);
                for (uint32_t j = lane; j < qp; j += kWave) {
                    const uint2 r = qrng[j];
                    cur[j] = lower_bound_sid(A.psid, A.fence, r.x, r.y, wa);
                }
                wave_sync();
                p1 = p0;
                continue;
            }
            // a streamed item still moves the cursors of the tokens it did not reach
            if (streamed) {
                for (uint32_t j = jdone + lane; j < qp; j += kWave) cur[j] = gallop_post(A.post, cur[j], qrng[j].y, wb);
                wave_sync();
            }
            if (lane == 0) {
                uint32_t k = (uint32_t)atomicAdd(&A.ctr[C_OVERFLOW], 1ull);
                if (k < A.ovf_cap) {
                    A.ovf[k] = (uint32_t)it;
                    A.ovfp[k] = streamed ? p0 : kEmpty;
                }
            }
            continue;
        }
        PHASE_MARK(1);

        // table slot accessors of phases C and D
        auto src_of = [&](uint32_t h) -> uint32_t { return tkey[h]; };
        auto cnt_of = [&](uint32_t h) -> uint32_t { return tcnt[h]; };
        auto last_of = [&](uint32_t h) -> uint32_t { return tlast[h]; };   // query position << 16 | source position
        // empties slot h when drop is set (the other lanes write a word of their own in the
        // idle query stage: no branch on a per-lane condition)
        auto drop_slot = [&](bool drop, uint32_t h) {
            *(drop ? &tkey[h] : &qstage[lane]) = kEmpty;
            *(drop ? &tcnt[h] : &qstage[lane]) = 0;
            *(drop ? &tlast[h] : &qstage[lane]) = 0;
        };
        uint32_t cand = 0, nsig = 0, nver = 0, inter = 0, vtok = 0, nres = 0;

        // ---- C: signature bound over clist[0, nc), one lane per candidate: 64 buckets,
        // then 128 for the rest; the survivors are compacted to the front of the list and
        // the others' slots emptied.  Returns the survivors.
        // (the source's bad-signature flag and its 64-bucket line are read together)
        auto phase_c = [&](const uint32_t nsig) -> uint32_t {
            uint32_t nver = 0, nex = 0;   // survivors; candidates that reach the 128-bucket bound
            // the query's counts are read through the scalar cache (wave-uniform operands of
            // the byte SADs), not gathered by every lane
            typedef __attribute__((address_space(4))) const uint32_t cu32q;
            const cu32q *qsp = (const cu32q *)(A.qsig + (size_t)q * kSigWords);
            const cu32q *qhp = (const cu32q *)(A.qsig64 + (size_t)q * kSig64Words);
            const bool qbad = A.qsigbad[q] != 0;
            // kSigPer candidates per lane and step: their signature lines are gathered
            // together, so a step waits for memory once
            constexpr int kSigPer = 2;
            for (uint32_t base = 0; base < nsig; base += kWave * kSigPer) {
                uint32_t ent[kSigPer], s[kSigPer];
                uint8_t sbad[kSigPer];
                uint4 b4[kSigPer][kSig64Words / 4];
#pragma unroll
                for (int u = 0; u < kSigPer; u++) {
                    const uint32_t c = base + (uint32_t)(u * kWave) + lane;
                    ent[u] = clist[min(c, nsig - 1u)];
                    s[u] = src_of(ent[u] & 0xFFFFu);
                }
#pragma unroll
                for (int u = 0; u < kSigPer; u++) {
                    sbad[u] = A.sigbad[s[u]];
                    const uint4 *sh4 = (const uint4 *)(A.sig64 + (size_t)s[u] * kSig64Words);
#pragma unroll
                    for (int k = 0; k < kSig64Words / 4; k++) b4[u][k] = sh4[k];
                }
                bool keep[kSigPer];
#pragma unroll
                for (int u = 0; u < kSigPer; u++) {
                    const bool in = base + (uint32_t)(u * kWave) + lane < nsig;
                    const uint32_t ls = ent[u] >> 16;
                    const int floorb = (int)floorf(fgss_min_overlap(lq, ls, t));
                    uint32_t sad = 0;
#pragma unroll
                    for (int k = 0; k < kSig64Words / 4; k++) {
                        sad = __builtin_amdgcn_sad_u8(b4[u][k].x, qhp[4 * k], sad);
                        sad = __builtin_amdgcn_sad_u8(b4[u][k].y, qhp[4 * k + 1], sad);
                        sad = __builtin_amdgcn_sad_u8(b4[u][k].z, qhp[4 * k + 2], sad);
                        sad = __builtin_amdgcn_sad_u8(b4[u][k].w, qhp[4 * k + 3], sad);
                    }
                    const bool bad = qbad || sbad[u] != 0;
                    // sum over buckets of min(cQ, cS) = (|Q| + |S| - SAD) / 2 >= |Q n S|
                    const bool exact = in && !bad && ((int)(lq + ls) - (int)sad) / 2 >= floorb;
                    keep[u] = in && bad;
                    const uint64_t exm = __ballot(exact);
                    nex += (uint32_t)__popcll(exm);
                    if (exm) {
                        if (exact) {
                            const uint4 *ss4 = (const uint4 *)(A.sig + (size_t)s[u] * kSigWords);
                            uint32_t sad2 = 0;
#pragma unroll
                            for (int k = 0; k < kSigWords / 4; k++) {
                                const uint4 c4 = ss4[k];
                                sad2 = __builtin_amdgcn_sad_u8(c4.x, qsp[4 * k], sad2);
                                sad2 = __builtin_amdgcn_sad_u8(c4.y, qsp[4 * k + 1], sad2);
                                sad2 = __builtin_amdgcn_sad_u8(c4.z, qsp[4 * k + 2], sad2);
                                sad2 = __builtin_amdgcn_sad_u8(c4.w, qsp[4 * k + 3], sad2);
                            }
                            keep[u] = ((int)(lq + ls) - (int)sad2) / 2 >= floorb;
                        }
                    }
                }
                // every entry of this step is in registers: compact in place (the
                // writes land at or below the entries just read) and clear the dropped
#pragma unroll
                for (int u = 0; u < kSigPer; u++) {
                    const bool in = base + (uint32_t)(u * kWave) + lane < nsig;
                    const uint32_t h = ent[u] & 0xFFFFu;
                    drop_slot(in && !keep[u], h);
                }
                wave_sync();
#pragma unroll
                for (int u = 0; u < kSigPer; u++) {
                    const uint64_t bal = __ballot(keep[u]);
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    *(keep[u] ? &clist[nver + below] : &qstage[lane]) = ent[u];
                    nver += (uint32_t)__popcll(bal);
                }
            }
            WACC(W_SIG2, nex);
            wave_sync();
            return nver;
        };

        // ---- D: verification (checking.cu:93-175) of clist[0, nver): overlap of S after
        // the last prefix match; the slots are emptied afterwards ----
        auto phase_d = [&](const uint32_t nver) {
            const bool staged = lq <= (uint32_t)kQStage;
            if (nver > 0 && staged) {
                for (uint32_t i = lane; i < lq; i += kWave) qstage[i] = A.qtok[qpos + i];
                wave_sync();
            }
            const uint32_t *Qv = staged ? qstage : A.qtok + qpos;
            for (uint32_t c = 0; c < nver; c++) {
                const uint32_t ent = clist[c];
                const uint32_t h = ent & 0xFFFFu, ls = ent >> 16;
                const uint32_t s = src_of(h);
                const uint32_t m = cnt_of(h);
                const uint32_t lastm = last_of(h);
                const uint32_t js = lastm >> 16, ss = lastm & 0xFFFFu;
                const uint32_t spos = A.pos[s];
                const float bound = fgss_min_overlap(lq, ls, t);
                const int floorb = (int)floorf(bound);
                uint32_t ov = m;
                uint32_t base = ss + 1;
                for (; base < ls; base += kWave) {
                    const uint32_t i = base + lane;
                    bool hit = false;
                    if (i < ls) {
                        // tokens of S after position ss exceed Q[js]: search Q(js, lq)
                        const uint32_t tk = A.tok[spos + i];
                        uint32_t lo = js + 1, hi = lq;
                        while (lo < hi) {
                            uint32_t mid = (lo + hi) >> 1;
                            if (Qv[mid] < tk) lo = mid + 1;
                            else hi = mid;
                        }
                        hit = lo < lq && Qv[lo] == tk;
                    }
                    ov += (uint32_t)__popcll(__ballot(hit));
                    const int rem = (int)ls - (int)(base + kWave);
                    if ((int)ov + (rem > 0 ? rem : 0) < floorb) { base += kWave; break; }
                }
                vtok += min(base, ls) - (ss + 1);
                if ((int)ov >= floorb) {
                    inter++;
                    if (fgss_similarity(lq, ls, ov) >= t) {            // block.cu:189-192
                        nres++;
                        if (lane == 0) {
                            uint32_t r = (uint32_t)atomicAdd(&A.ctr[C_RESULTS], 1ull);
                            if (r < A.out_cap) {
                                A.out[3 * (size_t)r] = (int32_t)q;
                                A.out[3 * (size_t)r + 1] = (int32_t)s;
                                A.out[3 * (size_t)r + 2] = (int32_t)ov;
                            }
                        }
                    }
                }
            }
            wave_sync();
            for (uint32_t c = lane; c < nver; c += kWave) drop_slot(true, clist[c] & 0xFFFFu);
            wave_sync();
        };

        {
            // ---- B: positional filter over the distinct sources ----
            // (branch-free like phase A: lanes past the list work on its last entry and
            // aim their stores at a word of their own in the idle query stage)
            uint32_t ncand = 0;
            for (uint32_t base = 0; base < nlist; base += kWave) {
                const uint32_t k = base + lane;
                const bool in = k < nlist;
                const uint32_t ent = clist[min(k, nlist - 1u)];
                const uint32_t h = ent & 0xFFFFu;
                const uint32_t ls = ent >> 16;
                const uint32_t m = tcnt[h];
                const uint32_t last = tlast[h];
                const float bound = fgss_min_overlap(lq, ls, t);
                const uint32_t js = last >> 16, ss = last & 0xFFFFu;
                const uint32_t rq = lq - js, rs = ls - ss;
                const uint32_t minrem = rq < rs ? rq : rs;
                const bool pass = in && !((float)(m + minrem) < bound);   // filtering.cu:127-133 at the last match
                cand += (uint32_t)__popcll(__ballot(pass));
                // overlap can still reach floor(bound): m + min(|Q|-js-1, |S|-ss-1)
                const bool keep = pass && (int)(m + minrem - 1) >= (int)floorf(bound);
                const bool drop = in && !keep;
                *(drop ? &tkey[h] : &qstage[lane]) = kEmpty;
                *(drop ? &tcnt[h] : &qstage[lane]) = 0;
                *(drop ? &tlast[h] : &qstage[lane]) = 0;
                const uint64_t bal = __ballot(keep);
                wave_sync();
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                *(keep ? &clist[ncand + below] : &qstage[lane]) = ent;
                ncand += (uint32_t)__popcll(bal);
            }
            wave_sync();
            wave_sync();
            PHASE_MARK(2);
            nsig = ncand;
            nver = phase_c(ncand);
            PHASE_MARK(3);
            phase_d(nver);
            PHASE_MARK(4);
        }
        WACC(W_JOB0 + 0, escan);
        WACC(W_JOB0 + 1, cand);
        WACC(W_JOB0 + 2, inter);
        WACC(W_JOB0 + 3, nres);
        WACC(W_ESCAN, escan);
        WACC(W_CAND, cand);
        WACC(W_SIG, nsig);
        WACC(W_VER, nver);
        WACC(W_INTER, inter);
        WACC(W_VTOK, vtok);
        if (nlist * 100 <= tsize * kSpanGrowPct) span = min(span * 2, kMaxSpan);
        else if (nlist * 100 > tsize * kSpanShrinkPct) span = max(span / 2, 1u);
        }   // windows
    }
    // this wavefront's totals and its last job's pending counters to its own slot:
    // one coalesced store, no atomics (thousands of wavefronts adding to the same
    // counters at the end of the launch serialize); k_wave_reduce sums the slots
    wave_sync();
    unsigned long long *slot = A.wctr + ((size_t)sblock * join_waves(CAP) + (size_t)w) * kWaveSlot;
    if (lane < kAccSlots) slot[lane] = wv;
    if (lane == kAccSlots) slot[lane] = cur_job;
}

// ---- Dense units (k_join_dense) ----
// A dense unit joins one dense query against one part of the global grid of DW = 2 x CAP
// records.  Its table is indexed by the source's offset in the grid part (record mod DW):
// dcnt holds 8-bit match counts, four slots to a word, and dlast the last match as query
// position << 24 | source position << 16 (both below 256: the probes of dense queries
// hold at most 255 tokens, and a source's mid-prefix position lies inside its partner's
// probe) | the low 16 bits of the source's postings, which hold its length (the same in
// every posting of a source, so the maximum keeps them).  Both tables are zero between
// units: the positional filter clears every word it reads.  A wavefront's LDS: the count
// tables of the block's wavefronts first (wave w at byte w x DW, so that a slot's count
// word is an OR of its offset and that base), then the last-match tables, then each
// wavefront's own area: its candidate list (two words an entry) -- which during the walk,
// while the list is empty, holds the probe-token table instead -- a word per lane for
// updates without work and a step's start marks.  Sharing the list's words with the token table keeps a block
// of four wavefronts at 512 table slots under 32 KiB, so that five blocks (as many as the
// kernel's registers allow) fit a CU's 160 KiB of LDS.
constexpr uint32_t kDList = 256;                          // candidates listed at once
constexpr int kDRows = 6;                                 // rows of kWave postings per full walk step
#ifndef FGSS_DENSE_SLOTS
#define FGSS_DENSE_SLOTS 4
#endif
constexpr uint32_t kDSlots = FGSS_DENSE_SLOTS;            // slots per lane in the positional filter (2 or 4)
static_assert(kDSlots == 2 || kDSlots == 4, "a lane's count bytes lie in one count word");
constexpr uint32_t kDMarkWords = kDRows * kWave / 4 + 1;  // start marks of a step, a byte per position, and a spare word
constexpr uint32_t kDOwnWords = 2 * kDList + kWave + kDMarkWords + 1;
static_assert(kDOwnWords % 2 == 0 && 2 * kDList >= 2 * kStreamQp, "entries stay 8-byte aligned; the token table fits the list");
constexpr uint32_t kDBad = 0x100u;                        // list entry: the source's signature overflowed
constexpr uint32_t dense_block_words(int cap) { return (uint32_t)join_waves(cap) * (2u * cap / 4u + 2u * cap + kDOwnWords); }
#ifndef FGSS_DENSE_LANES
#define FGSS_DENSE_LANES 2
#endif
constexpr int kDLanes = FGSS_DENSE_LANES;   // lanes per candidate in the 64-bucket signature test (1, 2 or 4)
static_assert(kDLanes == 1 || kDLanes == 2 || kDLanes == 4, "a candidate's lanes share a quad");

template <int CAP>
__device__ __forceinline__ void dense_units(const JoinArgs &A, uint32_t rblock, uint32_t rblocks)
{
    constexpr uint32_t DW = 2u * CAP;
    constexpr int LOG2DW = CAP == 512 ? 10 : CAP == 1024 ? 11 : CAP == 2048 ? 12 : 13;
    static_assert((1u << LOG2DW) == DW && DW <= 0xFFFFu, "grid parts of a power of two records, slots in 16 bits");
    static_assert(DW <= kSigbadPad && DW % (16u * kWave) == 0, "a grid part's overflow bytes are read 16 per lane");
    constexpr int LSH = 24 - LOG2DW;   // a dense posting: position << 24 | slot << LSH | length
    static_assert(LSH >= 11 && LSH <= 16, "the length field holds 2047 tokens or more, inside a last-match word's low 16 bits");
    constexpr int NW = join_waves(CAP);
    constexpr int PER = kStreamQp / kWave;   // probe tokens per lane
    extern __shared__ uint32_t smem[];
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = lane_id();
    const uint32_t cbase = w * DW;   // byte address of the wave's count words (a multiple of DW)
    uint32_t *const dlast = smem + (size_t)NW * (DW / 4u) + (size_t)w * DW;
    uint32_t *const own = smem + (size_t)NW * (DW / 4u + DW) + (size_t)w * kDOwnWords;
    uint2 *const list = (uint2 *)own;
    uint32_t *const cur4 = own;                // walk only: per non-empty probe token, in token order: 4 x (first posting - flattened start)
    uint32_t *const jw = cur4 + kStreamQp;     // ... and its query position
    uint32_t *const dump = own + 2 * kDList;   // a word per lane for updates without work
    uint32_t *const mw = dump + kWave;         // a step's start marks (words, for clearing)
    uint8_t *const smark = (uint8_t *)mw;
    auto cnt_word = [&](uint32_t h) -> uint32_t * {   // the count word of slot h
        return (uint32_t *)((char *)smem + ((h & (DW - 4u)) | cbase));
    };

    const uint64_t n_items = A.n_tiles_dev == nullptr || *A.n_items_dev > A.n_items_cap ? 0ull : *A.n_tiles_dev * kDenseTile;
    const uint32_t ndq = n_items ? *A.ndq_dev : 0u;
    unsigned long long *const slot = A.wctr + ((size_t)rblock * NW + w) * kWaveSlot;
    const uint64_t nwaves_grid = (uint64_t)rblocks * NW;
    const uint64_t g0 = min((n_items + nwaves_grid - 1) / nwaves_grid, (uint64_t)kQueueBatch);
    const uint64_t gw = (uint64_t)rblock * NW + w;
    if (gw * g0 >= n_items && nwaves_grid * g0 >= n_items) {   // nothing to claim
        if (lane < kAccSlots) slot[lane] = 0ull;
        if (lane == kAccSlots) slot[lane] = kEmpty;
        return;
    }
    for (uint32_t i = lane; i < DW / 4u; i += kWave) *cnt_word(i * 4u) = 0u;
    for (uint32_t i = lane; i < DW; i += kWave) dlast[i] = 0u;
    wave_sync();

    unsigned long long wv = 0;   // counter k in lane k (see join_items)
    unsigned long long c0 = clock64(), c1 = 0;
    (void)c0;
    (void)c1;
    auto lane_u64 = [&](int k) -> unsigned long long {   // counter k, wave-uniform
        return (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wv, k) |
               ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(wv >> 32), k) << 32);
    };
    // the current job's postings, candidates and intersections into the run totals (the
    // job counters hold every unit's, the totals none until the job's counters are left)
    auto fold_job = [&]() {
        const unsigned long long e = lane_u64(W_JOB0 + 0), cd = lane_u64(W_JOB0 + 1), in = lane_u64(W_JOB0 + 2);
        WACC(W_ESCAN, e);
        WACC(W_DPOST, e);
        WACC(W_CAND, cd);
        WACC(W_INTER, in);
    };
    const uint64_t qbase = nwaves_grid * g0;
    uint64_t batch_next = min(gw * g0, n_items), batch_end = min(gw * g0 + g0, n_items);
    uint32_t cur_job = kEmpty;
    float cur_t = 0.0f;
    uint32_t ugp = 0;
    uint64_t utlo = 1, uthi = 0, utile = ~0ull;
    uint32_t tq = kEmpty, thi = 0, tlen = 0, tpos = 0, tjob = 0;
    bool gclean = false;   // no source of the tile's grid part has an overflowed signature
    const uint32_t dng = (uint32_t)((A.n + DW - 1) / DW);
    // the next unit's probe tokens and offset-table entries, loaded ahead (pf_it: the unit)
    uint64_t pf_it = ~0ull;
    uint32_t pf_qp = 0, pf_tok[PER], pf_lo[PER], pf_hi[PER];
    for (;;) {
        if (batch_next >= batch_end) {
            if (qbase >= n_items) break;
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(&A.ctr[C_QUEUE_D], (unsigned long long)kQueueBatch);
            b0 = qbase + (((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(b0 >> 32)) << 32) |
                          (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)b0));
            batch_next = min((uint64_t)b0, n_items);
            batch_end = min((uint64_t)b0 + kQueueBatch, n_items);
        }
        if (batch_next >= n_items) break;
        const uint64_t it = batch_next++;
        // unit it: dense query it % kDenseTile of tile it / kDenseTile in part-major order
        // (k_dense_tiles); the grid part holding the tile is searched only when the tile
        // leaves the last part's range.  The tile's last units may lie past the part's
        // queries: their windows do not meet it.
        const uint64_t tile = it / kDenseTile;
        if (tile < utlo || tile >= uthi) {
            uint32_t lo = 0, hi = dng;   // the last part g with tpre[g] <= tile
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (A.dtpre[mid] <= tile) lo = mid;
                else hi = mid;
            }
            ugp = lo;
            utlo = A.dtpre[lo];
            uthi = A.dtpre[lo + 1];
        }
        if (tile != utile) {   // the tile's queries, one lane each, loaded together
            utile = tile;
            const uint64_t di = (uint64_t)A.dqlo[ugp] + (tile - utlo) * kDenseTile + (uint32_t)lane;
            tq = lane < (int)kDenseTile && di < ndq ? A.dl[di] : kEmpty;
            if (tq != kEmpty) {
                thi = A.hi[tq];
                tlen = A.len[tq];
                tpos = A.pos[tq];
                tjob = A.job[tq];
            }
            // the grid part's signature-overflow bytes, 16 per lane (the byte array is
            // padded to whole grid parts with zeros): when none is set, the signature test
            // reads no flag per candidate
            uint32_t anyb = 0;
#pragma unroll
            for (uint32_t i = 0; i < DW / (16u * kWave); i++) {
                const uint4 b = ((const uint4 *)(A.sigbad + (size_t)ugp * DW))[i * kWave + (uint32_t)lane];
                anyb |= b.x | b.y | b.z | b.w;
            }
            gclean = __ballot(anyb != 0) == 0;
        }
        const uint32_t u = (uint32_t)(it % kDenseTile);
        const uint32_t q = __builtin_amdgcn_readlane(tq, u);
        if (q == kEmpty) continue;
        const uint32_t gbase = ugp * DW;   // the grid part's first record
        const uint32_t wa = max(q + 1u, gbase), wb = min(__builtin_amdgcn_readlane(thi, u), gbase + DW);
        if (wa >= wb) continue;   // the query's window does not meet this grid part
        const uint32_t jq = __builtin_amdgcn_readlane(tjob, u);
        const float t = jq == cur_job ? cur_t : A.tj[jq];
        const uint32_t lq = __builtin_amdgcn_readlane(tlen, u) & 0xFFFFu;
        const uint32_t qpos = __builtin_amdgcn_readlane(tpos, u);
        if (jq != cur_job) {   // per-job counters: flush when the job changes
            if (cur_job != kEmpty) {
#pragma unroll
                for (int k = 0; k < kJobCtr; k++) {
                    const unsigned long long v = lane_u64(W_JOB0 + k);
                    if (lane == 0 && v) atomicAdd(&A.ctr_job[(size_t)cur_job * kJobCtr + k], v);
                }
            }
            fold_job();
            wv = lane >= W_JOB0 && lane < W_JOB0 + kJobCtr ? 0ull : wv;
            cur_job = jq;
            cur_t = t;
        }
        const uint32_t qp = min(fgss_max_prefix(lq, t), lq);
        PHASE_MARK(0);

        // ---- A: the postings of the probe tokens in [wa, wb).  A token's postings in the
        // grid part are offs[t][g] .. offs[t][g + 1] (k_dense_offs); in the window's first
        // or last part they are narrowed to [wa, wb) by search.  The non-empty tokens are
        // ranked in token order and their postings flattened: a step marks where each
        // token starts, a row's marks become a ballot, and each lane's token is the number
        // of starts at or before its position (mbcnt), so no posting needs a test. ----
        PHASE_START(r0);
        const uint2 *const row = A.doffs + (size_t)ugp * A.dterms;
        // probe token k x kWave + lane is this lane's k-th (each load reads 64 consecutive
        // tokens); the flattened order below is (lane, k), and every posting carries its
        // token's query position, so no step needs the tokens in query order
        uint32_t lo[PER], dc[PER];
        {
            if (pf_it == it) {   // prefetched by the unit before
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    lo[k] = pf_lo[k];
                    dc[k] = pf_hi[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const uint32_t j = (uint32_t)k * kWave + (uint32_t)lane;
                    uint32_t a = 0, b = 0;
                    if (j < qp) {
                        const uint2 ab = row[A.qtok[qpos + j]];
                        a = ab.x;
                        b = ab.y;
                    }
                    lo[k] = a;
                    dc[k] = b;
                }
            }
            // the next unit of the tile (the same grid part): its probe tokens now, their
            // offset-table entries after the walk
            pf_it = ~0ull;
            if (u + 1u < kDenseTile && it + 1 < batch_end) {
                const uint32_t q2 = __builtin_amdgcn_readlane(tq, u + 1u);
                if (q2 != kEmpty) {
                    const uint32_t jq2 = __builtin_amdgcn_readlane(tjob, u + 1u);
                    const float t2 = jq2 == jq ? t : A.tj[jq2];
                    const uint32_t lq2 = __builtin_amdgcn_readlane(tlen, u + 1u) & 0xFFFFu;
                    const uint32_t qpos2 = __builtin_amdgcn_readlane(tpos, u + 1u);
                    pf_qp = min(fgss_max_prefix(lq2, t2), lq2);
#pragma unroll
                    for (int k = 0; k < PER; k++) {
                        const uint32_t j = (uint32_t)k * kWave + (uint32_t)lane;
                        pf_tok[k] = j < pf_qp ? A.qtok[qpos2 + j] : 0u;
                    }
                    pf_it = it + 1;
                }
            }
            if (wa > gbase || wb < gbase + DW) {   // the window's first or last part
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    if (dc[k] > lo[k]) {
                        uint32_t a, b;
                        lower_bound_sid2(A.psid, A.fence, lo[k], dc[k], wa, lo[k], dc[k], wb, a, b);
                        lo[k] = a;
                        dc[k] = b;
                    }
                }
            }
        }
        uint32_t sdc = 0, nz = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            dc[k] -= lo[k];
            sdc += dc[k];
            nz += dc[k] != 0 ? 1u : 0u;
        }
        // one scan carries both the flattened starts (low 22 bits: at most 255 x 8192
        // postings) and the ranks of the non-empty tokens
        const uint32_t pk = sdc | nz << 22;
        const uint32_t inc = wave_inclusive_scan_dpp(pk);
        uint32_t run = (inc - pk) & 0x3FFFFFu, rk = (inc - pk) >> 22;
        uint32_t dbeg[PER];   // flattened start of each token (~0: empty)
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const bool ne = dc[k] != 0;
            *(ne ? &cur4[rk] : &dump[lane]) = (lo[k] - run) * 4u;
            *(ne ? &jw[rk] : &dump[lane]) = (uint32_t)k * kWave + (uint32_t)lane;
            dbeg[k] = ne ? run : ~0u;
            rk += ne ? 1u : 0u;
            run += dc[k];
        }
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1) & 0x3FFFFFu;
        PHASE_ADD(5, r0);
        const char *const pdb = (const char *)A.pden;
        uint32_t rb = 0;   // non-empty tokens starting before the current row
        // a step of N rows from flattened position e0; CHECK: the last step, whose rows
        // past nrows are skipped and whose lanes past the end neither load nor update
        auto wstep = [&](auto rows, auto check, uint32_t e0, uint32_t nrows) {
            constexpr int N = decltype(rows)::value;
            constexpr bool CHECK = decltype(check)::value;
            wave_sync();
#pragma unroll
            for (uint32_t i = 0; i < (uint32_t)N * 16u; i += kWave)
                if (i + lane < (uint32_t)N * 16u) mw[i + lane] = 0u;
            wave_sync();
            // a start outside the step (before e0 its offset wraps above) marks the spare
            // byte after the step's marks
#pragma unroll
            for (int k = 0; k < PER; k++) smark[min(dbeg[k] - e0, (uint32_t)(N * kWave))] = 1u;
            wave_sync();
            uint32_t pw[N], jv[N];
#pragma unroll
            for (int r = 0; r < N; r++) {
                if (CHECK && (uint32_t)r >= nrows) break;
                const uint64_t M = __ballot(smark[r * kWave + lane] != 0);
                const uint64_t X = M >> 1;
                const uint32_t acc = rb + (uint32_t)(M & 1ull) - 1u;
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(X >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)X, acc));
                rb += (uint32_t)__popcll(M);
                const uint32_t c4 = cur4[rank];
                jv[r] = jw[rank];
                uint32_t p = e0 + (uint32_t)(r * kWave) + (uint32_t)lane;
                if (CHECK) p = min(p, total - 1u);
                pw[r] = *(const uint32_t *)(pdb + (c4 + 4u * p));
            }
#pragma unroll
            for (int r = 0; r < N; r++) {
                if (CHECK && (uint32_t)r >= nrows) break;
                const uint32_t h = __builtin_amdgcn_ubfe(pw[r], (uint32_t)LSH, (uint32_t)LOG2DW);
                const uint32_t add = 1u << ((h << 3) & 31u);   // the count byte of slot h in its word
                // j << 24 | pos << 16 | the posting's low 16 bits (the source's length and
                // slot bits, the same in every posting of the source)
                const uint32_t last = __builtin_amdgcn_perm(jv[r], pw[r], 0x04030100u);
                if (CHECK) {
                    const bool v = e0 + (uint32_t)(r * kWave) + (uint32_t)lane < total;
                    atomicAdd(v ? cnt_word(h) : &dump[lane], add);
                    atomicMax(v ? &dlast[h] : &dump[lane], last);
                } else {
                    atomicAdd(cnt_word(h), add);
                    atomicMax(&dlast[h], last);
                }
            }
        };
        uint32_t e0 = 0;
        for (; e0 + (uint32_t)(kDRows * kWave) <= total; e0 += kDRows * kWave)
            wstep(std::integral_constant<int, kDRows>(), std::false_type(), e0, 0u);
        if (e0 < total) {
            const uint32_t nrows = (total - e0 + kWave - 1u) / kWave;   // 1 .. kDRows
            if (nrows > 2u) wstep(std::integral_constant<int, kDRows>(), std::true_type(), e0, nrows);
            else wstep(std::integral_constant<int, 2>(), std::true_type(), e0, nrows);
        }
        if (pf_it == it + 1) {
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const bool in = (uint32_t)k * kWave + (uint32_t)lane < pf_qp;
                const uint2 ab = in ? row[pf_tok[k]] : make_uint2(0u, 0u);
                pf_lo[k] = ab.x;
                pf_hi[k] = ab.y;
            }
        }
        wave_sync();
        PHASE_MARK(1);

        uint32_t nsig = 0, nver = 0, inter = 0, vtok = 0, nres = 0, nl = 0;
        // ---- C, C2 and D over list[0, nl) ----
        auto flush = [&]() {
            wave_sync();
            PHASE_MARK(2);
            nsig += nl;
            // C: 64 buckets, kDLanes lanes per candidate: each reads its part of the
            // candidate's 64-byte line (adjacent lanes, adjacent bytes: one instruction
            // touches kWave / kDLanes lines, not kWave), sums its part of the byte SADs
            // against the query's counts, and the parts are added across the lanes (DPP).
            // The query's counts are read once through the scalar cache.  Survivors (and
            // sources whose counts overflowed, marked kDBad) are compacted to the front of
            // the list by the first lane of each candidate.
            constexpr int LPC = kDLanes, WPL = kSig64Words / LPC, CPR = kWave / LPC;
            typedef __attribute__((address_space(4))) const uint32_t cu32;
            const cu32 *qhp = (const cu32 *)(A.qsig64 + (size_t)q * kSig64Words);
            uint32_t qh[kSig64Words];
#pragma unroll
            for (int k = 0; k < kSig64Words; k++) qh[k] = qhp[k];
            const uint32_t sub = (uint32_t)lane % LPC;
            uint32_t qv[WPL];   // the query's words of this lane's part
#pragma unroll
            for (int k = 0; k < WPL; k++) {
                uint32_t v = qh[k];
#pragma unroll
                for (int j = 1; j < LPC; j++) v = sub == (uint32_t)j ? qh[j * WPL + k] : v;
                qv[k] = v;
            }
            const bool qbad = A.qsigbad[q] != 0;
            const bool lead = sub == 0;
            uint32_t nv = 0, nex = 0;
            for (uint32_t base = 0; base < nl; base += CPR) {
                const uint32_t c = base + (uint32_t)lane / LPC;
                const bool in = c < nl;
                uint2 ent = list[min(c, nl - 1u)];
                const uint32_t s = gbase + (ent.x & 0xFFFFu);
                const uint8_t sbad = gclean ? (uint8_t)0 : A.sigbad[s];
                const uint4 *sh4 = (const uint4 *)(A.sig64 + (size_t)s * kSig64Words) + sub * (WPL / 4);
                uint4 b4[WPL / 4];
#pragma unroll
                for (int k = 0; k < WPL / 4; k++) b4[k] = sh4[k];
                const uint32_t ls = ent.x >> 16;
                const int floorb = (int)floorf(fgss_min_overlap(lq, ls, t));
                uint32_t sad = 0;
#pragma unroll
                for (int k = 0; k < WPL / 4; k++) {
                    sad = __builtin_amdgcn_sad_u8(b4[k].x, qv[4 * k], sad);
                    sad = __builtin_amdgcn_sad_u8(b4[k].y, qv[4 * k + 1], sad);
                    sad = __builtin_amdgcn_sad_u8(b4[k].z, qv[4 * k + 2], sad);
                    sad = __builtin_amdgcn_sad_u8(b4[k].w, qv[4 * k + 3], sad);
                }
                if (LPC >= 2) sad += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)sad, 0xB1, 0xf, 0xf, false);   // quad_perm 1,0,3,2
                if (LPC >= 4) sad += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)sad, 0x4E, 0xf, 0xf, false);   // quad_perm 2,3,0,1
                const bool bad = qbad || sbad != 0;
                // sum over buckets of min(cQ, cS) = (|Q| + |S| - SAD) / 2 >= |Q n S|
                const bool ok = ((int)(lq + ls) - (int)sad) / 2 >= floorb;
                nex += (uint32_t)__popcll(__ballot(lead && in && !bad && ok));
                const bool surv = lead && in && (bad || ok);
                ent.y |= bad ? kDBad : 0u;
                // every entry of this step is in registers: compact in place (the writes
                // land at or below the entries just read)
                wave_sync();
                const uint64_t bal = __ballot(surv);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                if (surv) list[nv + below] = ent;
                nv += (uint32_t)__popcll(bal);
            }
            WACC(W_SIG2, nex);
            wave_sync();
            // C2: the survivors with good counts against the 128-bucket counts, two lanes per
            // candidate as in C: each lane reads one 64-byte half of the candidate's line and
            // takes the SADs of its words against both halves of the query's counts (scalar
            // registers), keeping the one of its half -- twice the SADs of one lane per
            // candidate, half the lines per load instruction
            uint32_t nv2 = 0;
            {
                typedef __attribute__((address_space(4))) const uint32_t cu32q;
                const cu32q *qsp = (const cu32q *)(A.qsig + (size_t)q * kSigWords);
                constexpr int HW = kSigWords / 2;   // words per half line
                const uint32_t sub2 = (uint32_t)lane & 1u;
                for (uint32_t base = 0; base < nv; base += kWave / 2) {
                    const uint32_t c = base + (uint32_t)lane / 2u;
                    const bool in = c < nv;
                    const uint2 ent = list[min(c, nv - 1u)];
                    const bool bad = (ent.y & kDBad) != 0;
                    bool keep = in;
                    if (__ballot(in && !bad) != 0) {
                        const uint32_t s = gbase + (ent.x & 0xFFFFu), ls = ent.x >> 16;
                        const uint4 *ss4 = (const uint4 *)(A.sig + (size_t)s * kSigWords) + sub2 * (HW / 4);
                        uint32_t sa = 0, sb = 0;
                        if (in && !bad) {
#pragma unroll
                            for (int k = 0; k < HW / 4; k++) {
                                const uint4 c4 = ss4[k];
                                sa = __builtin_amdgcn_sad_u8(c4.x, qsp[4 * k], sa);
                                sa = __builtin_amdgcn_sad_u8(c4.y, qsp[4 * k + 1], sa);
                                sa = __builtin_amdgcn_sad_u8(c4.z, qsp[4 * k + 2], sa);
                                sa = __builtin_amdgcn_sad_u8(c4.w, qsp[4 * k + 3], sa);
                                sb = __builtin_amdgcn_sad_u8(c4.x, qsp[HW + 4 * k], sb);
                                sb = __builtin_amdgcn_sad_u8(c4.y, qsp[HW + 4 * k + 1], sb);
                                sb = __builtin_amdgcn_sad_u8(c4.z, qsp[HW + 4 * k + 2], sb);
                                sb = __builtin_amdgcn_sad_u8(c4.w, qsp[HW + 4 * k + 3], sb);
                            }
                        }
                        uint32_t sad2 = sub2 ? sb : sa;
                        sad2 += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)sad2, 0xB1, 0xf, 0xf, false);   // quad_perm 1,0,3,2
                        if (in && !bad) keep = ((int)(lq + ls) - (int)sad2) / 2 >= (int)floorf(fgss_min_overlap(lq, ls, t));
                    }
                    keep = keep && sub2 == 0u;
                    wave_sync();
                    const uint64_t bal = __ballot(keep);
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    if (keep) list[nv2 + below] = ent;
                    nv2 += (uint32_t)__popcll(bal);
                }
            }
            wave_sync();
            nver += nv2;
            PHASE_MARK(3);
            // D: verification (checking.cu:93-175): the source tokens after the last
            // prefix match, each searched in the query after its last match
            const uint32_t *Qv = A.qtok + qpos;
            for (uint32_t c = 0; c < nv2; c++) {
                const uint2 ent = list[c];
                const uint32_t s = gbase + (ent.x & 0xFFFFu), ls = ent.x >> 16;
                const uint32_t m = ent.y & 0xFFu, js = ent.y >> 24, ss = (ent.y >> 16) & 0xFFu;
                const uint32_t spos = A.pos[s];
                const int floorb = (int)floorf(fgss_min_overlap(lq, ls, t));
                uint32_t ov = m;
                uint32_t base = ss + 1;
                for (; base < ls; base += kWave) {
                    const uint32_t i = base + lane;
                    bool hit = false;
                    if (i < ls) {
                        const uint32_t tk = A.tok[spos + i];
                        uint32_t a = js + 1, b = lq;
                        while (a < b) {
                            const uint32_t mid = (a + b) >> 1;
                            if (Qv[mid] < tk) a = mid + 1;
                            else b = mid;
                        }
                        hit = a < lq && Qv[a] == tk;
                    }
                    ov += (uint32_t)__popcll(__ballot(hit));
                    const int rem = (int)ls - (int)(base + kWave);
                    if ((int)ov + (rem > 0 ? rem : 0) < floorb) { base += kWave; break; }
                }
                vtok += min(base, ls) - (ss + 1);
                if ((int)ov >= floorb) {
                    inter++;
                    if (fgss_similarity(lq, ls, ov) >= t) {   // block.cu:189-192
                        nres++;
                        if (lane == 0) {
                            const uint32_t r = (uint32_t)atomicAdd(&A.ctr[C_RESULTS], 1ull);
                            if (r < A.out_cap) {
                                A.out[3 * (size_t)r] = (int32_t)q;
                                A.out[3 * (size_t)r + 1] = (int32_t)s;
                                A.out[3 * (size_t)r + 2] = (int32_t)ov;
                            }
                        }
                    }
                }
            }
            PHASE_MARK(4);
            nl = 0;
        };

        // ---- B: the positional filter (filtering.cu:127-133) over the unit's slots,
        // kDSlots per lane (a lane's count bytes share a count word, its last-match words
        // are adjacent): every word read is cleared; the candidates that can still reach
        // floor(bound) are listed with their source length, count and last match, in slot
        // order.  A step's entries are counted before they are written: a list that cannot
        // take them goes through C, C2 and D first. ----
        const uint32_t slo = wa - gbase, shi = wb - gbase;
        constexpr uint32_t SL = kDSlots, STEP = SL * kWave;
        const uint32_t cshs = ((uint32_t)lane & (4u / SL - 1u)) * 8u * SL;   // the lane's count bytes in its word
        uint32_t candl = 0;   // candidates counted per lane, summed after the scan
        for (uint32_t h0 = slo & ~(STEP - 1u); h0 < shi; h0 += STEP) {
            const uint32_t hp = h0 + SL * (uint32_t)lane;   // slots hp .. hp + SL - 1 (< DW)
            const uint32_t cs = SL == 4 ? *cnt_word(hp) : (*cnt_word(hp) >> cshs) & ((1u << (8 * SL)) - 1u);
            if (__ballot(cs != 0) == 0) continue;   // slots outside [wa, wb) are never touched
            uint32_t wlv[SL];
            if constexpr (SL == 4) {
                uint4 *const p4 = (uint4 *)&dlast[hp];
                const uint4 v = *p4;
                *p4 = make_uint4(0u, 0u, 0u, 0u);
                wlv[0] = v.x; wlv[1] = v.y; wlv[2] = v.z; wlv[3] = v.w;
                *cnt_word(hp) = 0u;
            } else {
                uint2 *const p2 = (uint2 *)&dlast[hp];
                const uint2 v = *p2;
                *p2 = make_uint2(0u, 0u);
                wlv[0] = v.x; wlv[1] = v.y;
                *(uint16_t *)((char *)cnt_word(hp) + (cshs >> 3)) = 0;
            }
            uint32_t kb = 0;
#pragma unroll
            for (int i = 0; i < (int)SL; i++) {
                const uint32_t m = (cs >> (8 * i)) & 0xFFu;
                const uint32_t wl = wlv[i];
                const uint32_t ls = wl & ((1u << LSH) - 1u);   // the source's length, from its postings
                const uint32_t js = wl >> 24, ss = (wl >> 16) & 0xFFu;
                const float bound = fgss_min_overlap(lq, ls, t);
                const uint32_t rq = lq - js, rs = ls - ss;
                const uint32_t x = m + (rq < rs ? rq : rs);
                const bool pass = m != 0 && !((float)x < bound);
                candl += pass ? 1u : 0u;
                // overlap can still reach floor(bound): m + min(|Q|-js-1, |S|-ss-1)
                const bool keep = pass && (int)(x - 1u) >= (int)floorf(bound);
                kb |= (keep ? 1u : 0u) << i;
            }
            const uint32_t nk = (uint32_t)__popc(kb);
            const uint32_t incl = wave_inclusive_scan_dpp(nk);
            const uint32_t add = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
            if (nl + add > kDList) flush();   // wave-uniform; the step's entries are still in registers
            uint32_t at = nl + incl - nk;
#pragma unroll
            for (int i = 0; i < (int)SL; i++) {
                if ((kb >> i) & 1u) {
                    const uint32_t wl = wlv[i];
                    list[at++] = make_uint2((hp + (uint32_t)i) | (wl & ((1u << LSH) - 1u)) << 16,
                                            ((cs >> (8 * i)) & 0xFFu) | (wl & 0xFFFF0000u));
                }
            }
            nl += add;
        }
        const uint32_t cand = (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive_scan_dpp(candl), kWave - 1);
        if (nl) flush();
        wave_sync();
        // (the totals of postings, candidates and intersections are the job counters'
        // sums: fold_job adds them when a job's counters are flushed and at the end)
        WACC(W_JOB0 + 0, total);
        WACC(W_JOB0 + 1, cand);
        WACC(W_JOB0 + 2, inter);
        WACC(W_JOB0 + 3, nres);
        WACC(W_SIG, nsig);
        WACC(W_VER, nver);
        WACC(W_VTOK, vtok);
        WACC(W_DENSE, 1);
    }
    fold_job();
    wave_sync();
    if (lane < kAccSlots) slot[lane] = wv;
    if (lane == kAccSlots) slot[lane] = cur_job;
}

// Two kernels on the join's stream, one per role, so that each keeps its registers for
// one kind of work (one kernel with both roles spilled the dense loop's registers to
// scratch): k_join_dense walks the dense units, then k_join the item list.  Their
// wavefronts write disjoint counter slots: k_join's blocks follow the dense kernel's
// A.dense_blocks blocks.
template <int CAP>
__global__ __launch_bounds__(join_waves(CAP) * kWave, 4) void k_join_dense(JoinArgs A)
{
    if constexpr (CAP >= 512) dense_units<CAP>(A, blockIdx.x, gridDim.x);
}
template <int CAP, bool RS>
__global__ __launch_bounds__(join_waves(CAP) * kWave, 4) void k_join(JoinArgs A)
{
    join_items<CAP, RS>(A, blockIdx.x, gridDim.x, blockIdx.x + A.dense_blocks);
}
// Sum the per-wavefront slots of a k_join launch into the global counters and the
// per-job counters: a thread per slot.  Global counters: a wavefront sum, then one
// atomic per wavefront.  Per-job counters: consecutive slots usually share a job,
// so a segmented sum over equal job ids leaves one atomic per run of slots.
__global__ __launch_bounds__(256) void k_wave_reduce(const unsigned long long *__restrict__ wctr, uint32_t nwaves,
                                                     unsigned long long *__restrict__ ctr,
                                                     unsigned long long *__restrict__ ctr_job)
{
    constexpr int kMain = W_JOB0;   // global counters per slot
    const int dst[kMain] = {C_ESCAN, C_CAND, C_SIGCHK, C_VERIFIED, C_SIG2, C_INTER, C_VTOK, C_RETRY, C_DENSE,
                            C_DPOST, C_CYC_SETUP, C_CYC_SETUP + 1, C_CYC_SETUP + 2, C_CYC_SETUP + 3, C_CYC_SETUP + 4,
                            C_CYC_RANGES};
    const uint32_t wv = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = wv < nwaves;
    const unsigned long long *slot = wctr + (size_t)(in ? wv : 0) * kWaveSlot;
    const int lane = lane_id();
#pragma unroll
    for (int k = 0; k < kMain; k++) {
        const unsigned long long tot = wave_sum(in ? slot[k] : 0ull);
        if (lane == 0 && tot) atomicAdd(&ctr[dst[k]], tot);
    }
    const uint32_t job = in ? (uint32_t)slot[kAccSlots] : kEmpty;
    const uint32_t prev = __shfl_up(job, 1, kWave), next = __shfl_down(job, 1, kWave);
    const bool head = lane == 0 || prev != job, tail = lane == kWave - 1 || next != job;
    // each lane's segment start: a max-scan of head positions
    uint32_t start = head ? (uint32_t)lane : 0u;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(start, d, kWave);
        if (lane >= d) start = max(start, o);
    }
#pragma unroll
    for (int c = 0; c < kJobCtr; c++) {
        const unsigned long long v = (in && job != kEmpty) ? slot[W_JOB0 + c] : 0ull;
        const unsigned long long incl = wave_inclusive_scan(v);
        const unsigned long long before = __shfl(incl, (int)(start > 0 ? start - 1 : 0), kWave);
        const unsigned long long seg = incl - (start > 0 ? before : 0ull);
        if (tail && job != kEmpty && seg) atomicAdd(&ctr_job[(size_t)job * kJobCtr + c], seg);
    }
}

template <int CAP>
static size_t join_smem_bytes()
{
    return (size_t)join_waves(CAP) * join_wave_words(CAP) * sizeof(uint32_t);
}

// persistent grid: as many blocks as fit on the device, each wave loops over items
// persistent grid: as many blocks as fit on the device, each wave loops over items;
// with dense units possible, twice that: the dense role's blocks, then the other's
template <int CAP, bool RS>
static int launch_join(fgss_device_sets *d, const JoinArgs &A0, bool dry, bool with_dense)
{
    static thread_local int cached[3] = {-1, -1, -1};   // blocks, dense-kernel blocks, device
    const size_t smem = join_smem_bytes<CAP>(), smem_d = (size_t)dense_block_words(CAP) * sizeof(uint32_t);
    if (cached[0] < 0 || cached[2] != d->device) {   // runtime queries: never inside a graph capture
        int per_cu = 0, per_cu_d = 0, cus = 0;
        HIPCK(hipFuncSetAttribute((const void *)k_join<CAP, RS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_join<CAP, RS>, join_waves(CAP) * kWave, smem));
        if (CAP >= 512 && !RS) {
            HIPCK(hipFuncSetAttribute((const void *)k_join_dense<CAP>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_d));
            HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_d, k_join_dense<CAP>, join_waves(CAP) * kWave, smem_d));
        }
        HIPCK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device));
        cached[0] = std::max(per_cu, 1) * cus;
        cached[1] = std::max(per_cu_d, 1) * cus;
        cached[2] = d->device;
    }
    if (dry) return FGSS_OK;
    JoinArgs A = A0;
    A.dense_blocks = with_dense && CAP >= 512 && !RS && A.n_tiles_dev != nullptr ? (uint32_t)cached[1] : 0u;
    const uint32_t blocks = (uint32_t)cached[0];
    const uint32_t nw = (blocks + A.dense_blocks) * join_waves(CAP);
    if ((uint64_t)nw * kWaveSlot * 8 > d->b_wctr.cap)
        return fgss_fail(FGSS_ERR_DEVICE, "internal: wavefront counter slots too small");
    if (A.dense_blocks) {
        hipLaunchKernelGGL((k_join_dense<CAP>), dim3(A.dense_blocks), dim3(join_waves(CAP) * kWave), smem_d, d->st, A);
        HIPCK(hipGetLastError());
    }
    hipLaunchKernelGGL((k_join<CAP, RS>), dim3(blocks), dim3(join_waves(CAP) * kWave), smem, d->st, A);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_wave_reduce, dim3((nw + 255) / 256), dim3(256), 0, d->st, (const unsigned long long *)A.wctr,
                       nw, A.ctr, A.ctr_job);
    HIPCK(hipGetLastError());
    return FGSS_OK;
}

// with_dense: the launch's items may be dense (re-split items never are)
static int launch_join_cap(fgss_device_sets *d, const JoinArgs &A, int cap, bool dry = false, bool rs = false,
                           bool with_dense = true)
{
    if (rs) {
        switch (cap) {
        case 256: return launch_join<256, true>(d, A, dry, false);
        case 512: return launch_join<512, true>(d, A, dry, false);
        case 1024: return launch_join<1024, true>(d, A, dry, false);
        case 2048: return launch_join<2048, true>(d, A, dry, false);
        case 4096: return launch_join<4096, true>(d, A, dry, false);
        default: break;
        }
    }
    switch (cap) {
    case 256: return launch_join<256, false>(d, A, dry, with_dense);
    case 512: return launch_join<512, false>(d, A, dry, with_dense);
    case 1024: return launch_join<1024, false>(d, A, dry, with_dense);
    case 2048: return launch_join<2048, false>(d, A, dry, with_dense);
    case 4096: return launch_join<4096, false>(d, A, dry, with_dense);
    default: return fgss_fail(FGSS_ERR_ARG, "table_slots must be 256, 512, 1024, 2048 or 4096");
    }
}

static inline unsigned grid_for(int64_t n, int threads) { return (unsigned)std::max<int64_t>((n + threads - 1) / threads, 1); }

// device totals written by the scans (indices into b_misc); M_NDP64 (two words) = dense
// parts of the run
// device-side counts; M_NB64 (two words, 8-byte aligned) = boundary-table rows wanted
enum Misc { M_NE = 0, M_NITEMS, M_NP, M_NRESPLIT, M_NDENSE, M_NB64 = 6, M_NDP64 = 8, M_NDU64 = 10, M_COUNT = 16 };
constexpr uint32_t kBoundsBudget = 1u << 28;   // boundary-table entries at most (1 GiB)
constexpr unsigned kBoundsGrid = 2048;          // k_bounds blocks (the table size is known on the device only)

static uint64_t misc_nb(const fgss_device_sets *d)
{
    return (uint64_t)d->h_misc[M_NB64] | ((uint64_t)d->h_misc[M_NB64 + 1] << 32);
}

#ifndef FGSS_LOAD_MAX
#define FGSS_LOAD_MAX 100
#endif
#ifndef FGSS_DENSE_PCT
#define FGSS_DENSE_PCT 25
#endif
struct RunCfg {
    int cap = 512;          // LDS table slots per wave
    uint32_t target = 332;  // postings per work item
    uint32_t heavy_target = 409;   // postings per work item of a query of more than kHeavyParts parts
    uint32_t bounds = 1u << 28;   // boundary-table entries at most
    uint32_t stream_parts = kStreamParts;   // queries split into more parts are streamed
    uint32_t dense_w = 1024;      // sources per dense part (2 x cap; 0: no dense parts)
    uint32_t dense_pct = FGSS_DENSE_PCT;   // dense parts for windows whose postings fill them this full (%)
    uint32_t block_records = 0;   // sources per work item at most (0: no limit)
    int groups = 1;               // batches: job groups, one pipeline run each
};

static RunCfg make_cfg(const fgss_join_opts *opts)
{
    RunCfg c;
    c.cap = opts && opts->table_slots ? opts->table_slots : 512;
    const int load = opts && opts->load_pct > 0 ? std::min(opts->load_pct, FGSS_LOAD_MAX) : 65;
    c.target = std::max<uint32_t>(1u, (uint32_t)(c.cap * load / 100));
    // heavy queries at 80% unless a load is asked for (1M long records at t = 0.8:
    // k_join 2.01 s with every part at 65%, 1.93 s with heavy queries at 80%; 100k
    // titles, whose items are small, are slower with every part at 80%)
    c.heavy_target = opts && opts->load_pct > 0 ? c.target : std::max<uint32_t>(1u, (uint32_t)(c.cap * 80 / 100));
    if (opts && opts->bounds_cap > 0) c.bounds = std::min<uint32_t>((uint32_t)opts->bounds_cap, c.bounds);
    if (opts && opts->stream_parts > 0) c.stream_parts = (uint32_t)opts->stream_parts;
    // dense parts need the list room left after the cursor records: 512 table slots or more
    c.dense_w = c.cap >= 512 && !(opts && opts->dense_pct < 0) ? 2u * (uint32_t)c.cap : 0u;
    if (opts && opts->dense_pct > 0) c.dense_pct = (uint32_t)std::min(opts->dense_pct, 1000);
    if (opts && opts->block_records > 0) c.block_records = (uint32_t)opts->block_records;
    if (opts && opts->index_groups > 1) c.groups = opts->index_groups;
    return c;
}

// Grow a pipeline buffer; any reallocation invalidates a captured graph.
static int grow(fgss_device_sets *d, DevBuf &b, size_t bytes)
{
    const void *before = b.p;
    int rc = ensure(d, b, bytes);
    if (rc) return rc;
    if (b.p != before) d->cap_gen++;
    return FGSS_OK;
}

#define GROW(buf, bytes)                      \
    do {                                      \
        int rc_ = grow(d, (buf), (bytes));    \
        if (rc_) return rc_;                  \
    } while (0)

// Slots for the per-wavefront counters of a k_join launch: kWaveSlot words for
// every wavefront the device can hold (32 per CU), so any persistent grid fits.
static int wave_slots(fgss_device_sets *d)
{
    int cus = 0;
    HIPCK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device));
    return grow(d, d->b_wctr, (size_t)std::max(cus, 1) * 2 * 32 * kWaveSlot * 8);   // the two roles of a join launch
}

// Every buffer the pipeline touches, sized for upper bounds where one exists (index
// entries and probe tokens <= tokens) and for the current capacities otherwise.
// dense parts need an offset table of terms x grid parts words (k_dense_offs) within
// kDenseTableBudget at the finest grid, record ids below 2^24 and posting byte offsets in
// 32 bits; a dense posting holds a record's length in 24 - log2(grid width) bits
static bool dense_ok(const fgss_device_sets *d) { return d->n < ((int64_t)1 << 24) && d->ntok + kWave < ((int64_t)1 << 30); }
static uint32_t log2u(uint32_t x) { uint32_t k = 0; while ((2u << k) <= x) k++; return k; }
constexpr uint32_t kDenseGridMin = 1024;                   // records per grid part at 512 table slots
constexpr int64_t kDenseTableBudget = (int64_t)2 << 30;   // bytes
static int64_t dense_grid_parts(const fgss_device_sets *d, uint32_t dw) { return (d->n + dw - 1) / dw; }
static bool dense_table_fits(const fgss_device_sets *d)
{
    return (int64_t)d->n_terms * dense_grid_parts(d, kDenseGridMin) * 8 <= kDenseTableBudget;
}
// the run's dense grid (records per grid part; 0: no dense parts)
static uint32_t dense_w_of(const fgss_device_sets *d, const RunCfg &cfg)
{
    if (cfg.dense_w == 0 || !dense_ok(d) || !dense_table_fits(d)) return 0u;
    return (int64_t)d->max_len < ((int64_t)1 << (24 - log2u(cfg.dense_w))) ? cfg.dense_w : 0u;
}


static int prepare_buffers(fgss_device_sets *d)
{
    const int64_t n = std::max<int64_t>(d->n, 1), E = std::max<int64_t>(d->ntok, 1);
    if (d->cap_items == 0) d->cap_items = (uint32_t)std::min<int64_t>(2 * n + E / 64 + 4096, UINT32_MAX / 2);
    if (d->cap_bounds == 0) d->cap_bounds = (uint32_t)std::min<int64_t>(E + 4096, kBoundsBudget);
    if (d->cap_out == 0) d->cap_out = 65536;
    GROW(d->b_midlen, (size_t)n * 4);
    GROW(d->b_midoff, (size_t)(n + 1) * 4);
    GROW(d->b_probelen, (size_t)n * 4);
    GROW(d->b_poff, (size_t)(n + 1) * 4);
    GROW(d->b_scanscr, (size_t)scan_scratch_elems(n + 1) * 8);   // also 64-bit scans
    GROW(d->b_ka, (size_t)E * 4);
    GROW(d->b_kb, (size_t)E * 4);
    GROW(d->b_ia, (size_t)E * 4);
    GROW(d->b_ib, (size_t)E * 4);
    GROW(d->b_sid_of, (size_t)E * 4);
    GROW(d->b_pl_of, (size_t)E * 4);
    // kWave spare postings: a streamed round reads up to 63 postings past the end of a
    // token's list before it tests them against the end
    GROW(d->b_post, (size_t)(E + kWave) * 8);
    if (dense_ok(d)) GROW(d->b_pden, (size_t)(E + kWave) * 4);
    if (dense_ok(d) && dense_table_fits(d)) {   // sized for the finest grid (the largest table)
        const int64_t ng = dense_grid_parts(d, kDenseGridMin);
        GROW(d->b_doffs, std::max<size_t>((size_t)d->n_terms * (size_t)ng * 8, 16));
        GROW(d->b_dq, (size_t)(ng + 1) * 4);
        GROW(d->b_dtpre, (size_t)(ng + 1) * 8);
    }
    GROW(d->b_psid, (size_t)(E + 1) * 4);
    GROW(d->b_fence, (size_t)(E / kFence + 2) * 4);
    GROW(d->b_loff, ((size_t)d->n_terms + 2) * 4);
    GROW(d->b_sortscr, (size_t)radix_scratch_elems(E) * 4);
    GROW(d->b_rng, (size_t)E * sizeof(uint2));
    GROW(d->b_hi, (size_t)n * 4);
    GROW(d->b_eq, (size_t)n * 4);
    GROW(d->b_nparts, (size_t)n * 4);
    GROW(d->b_partoff, (size_t)(n + 1) * 4);
    GROW(d->b_bcnt, (size_t)n * 8);
    GROW(d->b_bofs, (size_t)(n + 1) * 8);
    GROW(d->b_nitems, (size_t)n * 4);
    GROW(d->b_dflag, (size_t)n * 4);
    GROW(d->b_doff, (size_t)(n + 1) * 4);
    GROW(d->b_bounds, (size_t)d->cap_bounds * 4);
    GROW(d->b_items, (size_t)d->cap_items * sizeof(Item));
    GROW(d->b_dlist, (size_t)n * 4);   // the dense query list
    GROW(d->b_grp, 8);
    d->cap_ovf = std::max(d->cap_ovf, d->cap_items);
    GROW(d->b_ovf, (size_t)d->cap_ovf * 4);
    GROW(d->b_ovfp, (size_t)d->cap_ovf * 4);
    {
        int rc_ = wave_slots(d);
        if (rc_) return rc_;
    }
    GROW(d->b_out, (size_t)d->cap_out * 12);
    GROW(d->b_ctr, C_COUNT * sizeof(uint64_t));
    GROW(d->b_jctr, (size_t)d->njobs * kJobCtr * sizeof(uint64_t));
    GROW(d->b_misc, M_COUNT * 4);
    return FGSS_OK;
}

static JoinArgs join_args(fgss_device_sets *d)
{
    JoinArgs A;
    A.tok = d->tok; A.pos = d->pos; A.len = d->len;
    A.loff = P<uint32_t>(d->b_loff); A.post = P<uint2>(d->b_post);
    A.psid = P<uint32_t>(d->b_psid); A.fence = P<uint32_t>(d->b_fence);
    A.sig = d->sig; A.sig64 = d->sig64; A.sigbad = d->sigbad;
    A.qtok = d->tok; A.qsig = d->sig; A.qsig64 = d->sig64; A.qsigbad = d->sigbad;
    A.rng = P<uint2>(d->b_rng); A.poff = P<uint32_t>(d->b_poff); A.bounds = P<uint32_t>(d->b_bounds);
    A.pden = dense_ok(d) ? P<uint32_t>(d->b_pden) : nullptr;
    A.doffs = P<uint2>(d->b_doffs); A.dterms = (uint32_t)d->n_terms; A.dqlo = P<uint32_t>(d->b_dq); A.dtpre = P<uint64_t>(d->b_dtpre);
    A.dl = P<uint32_t>(d->b_dlist); A.ndq_dev = P<uint32_t>(d->b_misc) + M_NDENSE;
    A.n_tiles_dev = d->b_doffs.p ? (const unsigned long long *)(P<uint32_t>(d->b_misc) + M_NDU64) : nullptr;
    A.hi = P<uint32_t>(d->b_hi); A.nparts = P<uint32_t>(d->b_nparts); A.n = d->n;
    A.items = P<Item>(d->b_items);
    A.n_items_dev = P<uint32_t>(d->b_misc) + M_NITEMS; A.n_items_cap = d->cap_items;
    A.dense_blocks = 0;
    A.job = d->job; A.tj = d->tj; A.ctr_job = P<unsigned long long>(d->b_jctr);
    A.out = P<int32_t>(d->b_out); A.out_cap = d->cap_out; A.ovf = P<uint32_t>(d->b_ovf);
    A.ovfp = P<uint32_t>(d->b_ovfp); A.ovf_cap = d->cap_ovf;
    A.wctr = P<unsigned long long>(d->b_wctr);
    A.ctr = P<unsigned long long>(d->b_ctr);
    return A;
}

// Index build (steps 1-5) on d->st: per record the indexed prefix (self-join:
// mid_prefix at the record's job threshold; rs: max_prefix at tj[0], the smallest
// threshold an R-S index will serve) and the probe prefix, the stable sort of the
// entries by term, packed postings, the id-only copy, the fence and list offsets.
// dense_w: the run's dense grid width (0: no dense postings)
static int enqueue_index(fgss_device_sets *d, int rs, uint32_t dense_w)
{
    hipStream_t st = d->st;
    const int64_t n = d->n, E = std::max<int64_t>(d->ntok, 1);
    uint32_t *misc = P<uint32_t>(d->b_misc), *scr = P<uint32_t>(d->b_scanscr);
    uint32_t *midlen = P<uint32_t>(d->b_midlen), *midoff = P<uint32_t>(d->b_midoff);
    uint32_t *probelen = P<uint32_t>(d->b_probelen), *poff = P<uint32_t>(d->b_poff);
    uint32_t *ka = P<uint32_t>(d->b_ka), *kb = P<uint32_t>(d->b_kb), *ia = P<uint32_t>(d->b_ia),
             *ib = P<uint32_t>(d->b_ib), *sid_of = P<uint32_t>(d->b_sid_of), *loff = P<uint32_t>(d->b_loff);
    uint32_t *psid = P<uint32_t>(d->b_psid), *fence = P<uint32_t>(d->b_fence);
    uint2 *post = P<uint2>(d->b_post);
    hipLaunchKernelGGL(k_midlen, dim3(grid_for(n, 256)), dim3(256), 0, st, d->len, d->job, d->tj, midlen, probelen, n, rs,
                       (int64_t)d->q_begin, (int64_t)d->q_end, rs ? (const uint32_t *)nullptr : P<uint32_t>(d->b_grp));
    HIPCK(hipGetLastError());
    HIPCK(exclusive_scan<uint32_t>(midlen, midoff, n, misc + M_NE, scr, st));
    HIPCK(exclusive_scan<uint32_t>(probelen, poff, n, misc + M_NP, scr, st));
    hipLaunchKernelGGL(k_entries, dim3(grid_for(n, 256)), dim3(256), 0, st, d->tok, d->pos, d->len, midoff, misc + M_NE,
                       n, ka, sid_of, P<uint32_t>(d->b_pl_of));
    HIPCK(hipGetLastError());
    int bits = 1;
    while (bits < 32 && (1ll << bits) < (long long)d->n_terms) bits++;
    int which = 0;
    HIPCK(radix_sort_pairs(ka, ia, kb, ib, E, misc + M_NE, bits, P<uint32_t>(d->b_sortscr), st, &which));
    const uint32_t *skeys = which ? kb : ka;
    const uint32_t *sidx = which ? ib : ia;
    d->skeys = skeys;
    hipLaunchKernelGGL(k_postings, dim3(grid_for(E, 256)), dim3(256), 0, st, sidx, sid_of, P<uint32_t>(d->b_pl_of),
                       misc + M_NE, post, psid, fence, rs || !dense_w ? (uint32_t *)nullptr : P<uint32_t>(d->b_pden),
                       dense_w, dense_w ? 24u - log2u(dense_w) : 0u);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_list_offsets, dim3(grid_for(E + 1, 256)), dim3(256), 0, st, skeys, misc + M_NE, d->n_terms, loff);
    HIPCK(hipGetLastError());
    return FGSS_OK;
}

// A run's counters (ctr), per-job counters (jctr) and small counters (misc) cleared
// in one launch
__global__ void k_clear_run(unsigned long long *__restrict__ ctr, int nctr, unsigned long long *__restrict__ jctr,
                            int64_t njctr, uint32_t *__restrict__ misc, int nmisc)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < njctr; i += (int64_t)gridDim.x * blockDim.x)
        jctr[i] = 0ull;
    if (blockIdx.x == 0) {
        for (int i = (int)threadIdx.x; i < nctr; i += (int)blockDim.x) ctr[i] = 0ull;
        for (int i = (int)threadIdx.x; i < nmisc; i += (int)blockDim.x) misc[i] = 0u;
    }
}

static hipError_t clear_run(fgss_device_sets *d, hipStream_t st)
{
    const int64_t nj = (int64_t)d->njobs * kJobCtr;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nj + 255) / 256, 4096));
    hipLaunchKernelGGL(k_clear_run, dim3(blocks), dim3(256), 0, st, P<unsigned long long>(d->b_ctr), (int)C_COUNT,
                       P<unsigned long long>(d->b_jctr), nj, P<uint32_t>(d->b_misc), (int)M_COUNT);
    return hipGetLastError();
}

// The self-join pipeline on d->st in three stages with no host synchronization (so
// that each can be captured into a HIP graph): the index build (counters cleared
// first), the planning of work items, and the join kernel; then the device->host
// copies of all counters into pinned memory.
static int enqueue_stage_index(fgss_device_sets *d, const RunCfg &cfg)
{
    HIPCK(clear_run(d, d->st));
    return enqueue_index(d, 0, dense_w_of(d, cfg));
}

static int enqueue_stage_plan(fgss_device_sets *d, const RunCfg &cfg)
{
    hipStream_t st = d->st;
    const int64_t n = d->n;
    uint32_t *misc = P<uint32_t>(d->b_misc), *scr = P<uint32_t>(d->b_scanscr);
    uint32_t *probelen = P<uint32_t>(d->b_probelen), *poff = P<uint32_t>(d->b_poff);
    uint32_t *loff = P<uint32_t>(d->b_loff), *psid = P<uint32_t>(d->b_psid), *fence = P<uint32_t>(d->b_fence);
    uint32_t *hi = P<uint32_t>(d->b_hi), *eq = P<uint32_t>(d->b_eq), *nparts = P<uint32_t>(d->b_nparts),
             *partoff = P<uint32_t>(d->b_partoff);
    uint64_t *bcnt = P<uint64_t>(d->b_bcnt), *bofs = P<uint64_t>(d->b_bofs), *scr64 = P<uint64_t>(d->b_scanscr);
    // 6-8: plan work items; with dense parts on, k_plan flags the dense queries
    const uint32_t dense_w = dense_w_of(d, cfg);
    const uint32_t ng = dense_w ? (uint32_t)dense_grid_parts(d, dense_w) : 0u;
    uint32_t *dflag = P<uint32_t>(d->b_dflag), *doff = P<uint32_t>(d->b_doff), *dl = P<uint32_t>(d->b_dlist);
    if (dense_w) {   // the dense offset table: k_plan counts the postings of dense candidates in it
        hipLaunchKernelGGL(k_dense_offs, dim3((unsigned)std::min<int64_t>(grid_for(std::max<int64_t>(d->ntok, 1), 256),
                                                                           kDenseOffsGrid)),
                           dim3(256), 0, st, d->skeys, psid, loff, misc + M_NE, d->n_terms, dense_w, ng,
                           P<uint32_t>(d->b_doffs));
        HIPCK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_plan, dim3((unsigned)std::min<int64_t>(grid_for(n * kPlanGroup, 256), 4096)), dim3(256), 0, st,
                       d->tok, d->pos, d->len, n, d->job, d->tj, d->jend, probelen, poff, loff, psid, fence, cfg.target,
                       cfg.heavy_target, P<uint2>(d->b_rng), hi, nparts, eq, bcnt, cfg.stream_parts,
                       dense_w, cfg.dense_pct, (unsigned long long *)(misc + M_NDP64), cfg.block_records,
                       dense_w ? dflag : (uint32_t *)nullptr, dense_w ? (const uint2 *)P<uint2>(d->b_doffs) : nullptr,
                       (uint32_t)d->n_terms);
    HIPCK(hipGetLastError());
    if (dense_w) {   // the dense query list and work units (nothing when none were planned)
        const unsigned long long *ndp = (const unsigned long long *)(misc + M_NDP64);
        HIPCK(exclusive_scan<uint32_t>(dflag, doff, n, misc + M_NDENSE, scr, st));
        hipLaunchKernelGGL(k_dense_list, dim3(grid_for(n, 256)), dim3(256), 0, st, (const uint32_t *)dflag,
                           (const uint32_t *)doff, n, ndp, dl);
        HIPCK(hipGetLastError());
        uint64_t *tpre = P<uint64_t>(d->b_dtpre);
        hipLaunchKernelGGL(k_dense_tiles, dim3(grid_for((int64_t)ng + 1, 256)), dim3(256), 0, st, (const uint32_t *)dl,
                           misc + M_NDENSE, hi, ng, dense_w, ndp, P<uint32_t>(d->b_dq), tpre);
        HIPCK(hipGetLastError());
        HIPCK(exclusive_scan<uint64_t>(tpre, tpre, (int64_t)ng + 1, (uint64_t *)(misc + M_NDU64), scr64, st));
    }
    const uint32_t blim = std::min(d->cap_bounds, cfg.bounds);
    uint32_t *nitems = P<uint32_t>(d->b_nitems);
    HIPCK(exclusive_scan<uint64_t>(bcnt, bofs, n, (uint64_t *)(misc + M_NB64), scr64, st));
    // (dense queries have no items: the dense role of the join walks their units)
    hipLaunchKernelGGL(k_nitems, dim3(grid_for(n, 256)), dim3(256), 0, st, nparts, probelen, bofs, bcnt, blim, n, nitems);
    HIPCK(hipGetLastError());
    HIPCK(exclusive_scan<uint32_t>(nitems, partoff, n, misc + M_NITEMS, scr, st));
    hipLaunchKernelGGL(k_bounds, dim3(kBoundsGrid), dim3(256), 0, st, n, probelen, poff, P<uint2>(d->b_rng),
                       (const uint32_t *)nullptr, hi, nparts, bofs, bcnt, (const uint64_t *)(misc + M_NB64), psid,
                       fence, P<uint32_t>(d->b_bounds), blim);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_items, dim3(kItemsGrid), dim3(256), 0, st, 0, (const uint32_t *)nullptr, hi, nparts,
                       probelen, eq, partoff, bofs, bcnt, blim, d->job, d->tj,
                       d->len, d->pos, poff, n, misc + M_NITEMS, P<Item>(d->b_items), d->cap_items,
                       (const uint32_t *)d->jend);
    HIPCK(hipGetLastError());
    return FGSS_OK;
}

static int enqueue_stage_join(fgss_device_sets *d, const RunCfg &cfg)
{
    return launch_join_cap(d, join_args(d), cfg.cap, false, false, dense_w_of(d, cfg) != 0);   // 9: the join
}

// The first out_hint result triples into pinned memory, behind the join on its stream
// (the count is known on the device only; a run that finds more copies the rest
// after it completes).  Never inside a graph capture: it may allocate.
static int enqueue_out_copy(fgss_device_sets *d)
{
    const uint32_t want = std::min(d->cap_out, d->out_hint);
    d->out_copied = 0;
    if (want == 0) return FGSS_OK;
    if (d->h_out_cap < want) {   // pinned allocations are slow: grow by doubling at least
        const size_t cap = std::max<size_t>({(size_t)want, 2 * d->h_out_cap, 16384});
        retire_host(d, d->h_out);
        d->h_out = nullptr;
        d->h_out_cap = 0;
        HIPCK(hipHostMalloc((void **)&d->h_out, cap * 12, hipHostMallocDefault));
        d->h_out_cap = cap;
    }
    HIPCK(hipMemcpyAsync(d->h_out, d->b_out.p, (size_t)want * 12, hipMemcpyDeviceToHost, d->st));
    d->out_copied = want;
    return FGSS_OK;
}

static int enqueue_copies(fgss_device_sets *d)
{
    hipStream_t st = d->st;
    HIPCK(hipMemcpyAsync(d->h_ctr, d->b_ctr.p, C_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(d->h_misc, d->b_misc.p, M_COUNT * 4, hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(d->h_jctr, d->b_jctr.p, (size_t)d->njobs * kJobCtr * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         st));
    d->jctr_async = true;
    return enqueue_out_copy(d);
}

// The whole join, eager; timing events around the stages (ev[0], ev[1], ev[4], ev[5]).
static int enqueue_pipeline(fgss_device_sets *d, const RunCfg &cfg, bool timing)
{
    hipStream_t st = d->st;
    if (timing) HIPCK(hipEventRecord(d->ev[0], st));
    int rc = enqueue_stage_index(d, cfg);
    if (rc) return rc;
    if (timing) HIPCK(hipEventRecord(d->ev[1], st));
    rc = enqueue_stage_plan(d, cfg);
    if (rc) return rc;
    if (timing) HIPCK(hipEventRecord(d->ev[4], st));
    rc = enqueue_stage_join(d, cfg);
    if (rc) return rc;
    if (timing) HIPCK(hipEventRecord(d->ev[5], st));
    return enqueue_copies(d);
}

// FGSS_TRACE=1 prints a host-side timeline of each join to stderr
static void trace_mark(const char *what)
{
    static const bool on = getenv("FGSS_TRACE") && getenv("FGSS_TRACE")[0] == '1';
    static thread_local std::chrono::steady_clock::time_point last;
    if (!on) return;
    auto now = std::chrono::steady_clock::now();
    if (strcmp(what, "begin") != 0)
        fprintf(stderr, "[fgss] %-12s +%8.3f ms\n", what, std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
}

struct RunTotals {
    uint64_t acc[C_COUNT] = {0};   // global counters summed over launches
    std::vector<uint64_t> jctr;    // per job: kJobCtr counters
    uint32_t index_entries = 0, probe_tokens = 0;
    double ms_index = 0, ms_join = 0, ms_kernel = 0;
    int launches = 0;
    bool graph = false;
};

// Everything the captured stages bake into their launches besides device pointers
// (a reallocation of any buffer bumps cap_gen): the configuration, the capacities
// the kernels test against and the record counts.  A replay is valid while it is
// unchanged; a capacity can grow inside a buffer's slack without a reallocation.
static std::array<int64_t, 17> graph_key_of(const fgss_device_sets *d, const RunCfg &cfg)
{
    return {cfg.cap, cfg.target, cfg.heavy_target, cfg.bounds, cfg.stream_parts, d->cap_gen,
            d->cap_items, d->cap_out, d->cap_ovf, d->cap_bounds, d->n, d->njobs, d->q_begin, d->q_end,
            cfg.dense_w, cfg.dense_pct, cfg.block_records};
}

// For the length of one call: the record set's work on the caller's stream (opts->stream)
// and, for a self-join, its query range; both restored when the call returns.
struct CallScope {
    fgss_device_sets *d;
    hipStream_t own;
    int32_t qb, qe;
    CallScope(fgss_device_sets *dd, const fgss_join_opts *opts, bool ranged) : d(dd), own(dd->st), qb(dd->q_begin), qe(dd->q_end)
    {
        if (opts && opts->stream) d->st = (hipStream_t)opts->stream;
        d->q_begin = ranged && opts ? (int32_t)opts->query_begin : 0;
        d->q_end = ranged && opts ? (int32_t)opts->query_end : 0;
    }
    ~CallScope()
    {
        d->st = own;
        d->q_begin = qb;
        d->q_end = qe;
    }
};

// A query range in opts must lie inside the n records (0, 0 = all).
static int check_range(const fgss_join_opts *opts, int64_t n, bool allowed)
{
    if (!opts || (opts->query_begin == 0 && opts->query_end == 0)) return FGSS_OK;
    if (!allowed) return fgss_fail(FGSS_ERR_ARG, "a query range applies to plain self-joins only");
    if (opts->query_begin < 0 || opts->query_end < opts->query_begin || opts->query_end > n)
        return fgss_fail(FGSS_ERR_ARG, "query range [%lld, %lld) outside the %lld records",
                         (long long)opts->query_begin, (long long)opts->query_end, (long long)n);
    return FGSS_OK;
}

static int capture_graph(fgss_device_sets *d, const RunCfg &cfg)
{
    for (int k = 0; k < 3; k++) {
        if (d->graph_exec[k]) { HIPCK(hipGraphExecDestroy(d->graph_exec[k])); d->graph_exec[k] = nullptr; }
        if (d->graph[k]) { HIPCK(hipGraphDestroy(d->graph[k])); d->graph[k] = nullptr; }
    }
    for (int k = 0; k < 3; k++) {
        HIPCK(hipStreamBeginCapture(d->st, hipStreamCaptureModeThreadLocal));
        int rc = k == 0 ? enqueue_stage_index(d, cfg) : k == 1 ? enqueue_stage_plan(d, cfg) : enqueue_stage_join(d, cfg);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(d->st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fgss_fail(FGSS_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
        d->graph[k] = g;
        HIPCK(hipGraphInstantiate(&d->graph_exec[k], g, nullptr, nullptr, 0));
    }
    d->graph_key = graph_key_of(d, cfg);
    return FGSS_OK;
}

// How a record set runs its pipeline: a self-join (its own records are queries and
// sources) or an R-S join of a slot's probe records against a resident index.
struct PipeOps {
    std::function<int()> prepare;            // size every buffer (may allocate)
    std::function<int(bool)> enqueue;        // enqueue everything, no host syncs (arg: record timing events)
    std::function<JoinArgs()> args;          // join-kernel arguments (for re-split rounds)
    bool rs = false;
    bool graph_ok = false;
};

static PipeOps self_ops(fgss_device_sets *d, const RunCfg &cfg)
{
    PipeOps o;
    o.prepare = [d] { return prepare_buffers(d); };
    o.enqueue = [d, cfg](bool timing) { return enqueue_pipeline(d, cfg, timing); };
    o.args = [d] { return join_args(d); };
    o.graph_ok = true;
    return o;
}

// Enqueue one run (eager, or a replay of the captured graph).
static int launch_once(fgss_device_sets *d, const RunCfg &cfg, const PipeOps &ops, bool use_graph)
{
    int rc = ops.prepare();
    if (rc) return rc;
    if (!ops.rs) {   // the job group of this run, ahead of the (possibly captured) stages
        hipLaunchKernelGGL(k_set_pair, dim3(1), dim3(64), 0, d->st, P<uint32_t>(d->b_grp), d->grp_begin, d->grp_end);
        HIPCK(hipGetLastError());
    }
    if (use_graph && ops.graph_ok) {
        rc = launch_join_cap(d, ops.args(), cfg.cap, true, ops.rs);   // launch configuration, outside the capture
        if (rc) return rc;
        if (!d->graph_exec[0] || d->graph_key != graph_key_of(d, cfg)) {
            rc = capture_graph(d, cfg);
            if (rc) return rc;
        }
        // the stages' graphs with the timing events between them, then the copies
        const int evk[3] = {0, 1, 4};
        for (int k = 0; k < 3; k++) {
            HIPCK(hipEventRecord(d->ev[evk[k]], d->st));
            HIPCK(hipGraphLaunch(d->graph_exec[k], d->st));
        }
        HIPCK(hipEventRecord(d->ev[5], d->st));
        return enqueue_copies(d);
    }
    return ops.enqueue(true);
}

// After a launched run has completed (stream synchronized or its event done): rerun
// with grown buffers while any capacity was exceeded, then collect the counters and
// the (q, s, overlap) triples, then run host-driven rounds for the work items whose
// LDS table overflowed (each split four ways).
static int finish_run(fgss_device_sets *d, const RunCfg &cfg, const PipeOps &ops, bool use_graph,
                      std::vector<int32_t> &tri, RunTotals &T)
{
    hipStream_t st = d->st;
    tri.clear();
    T = RunTotals();
    T.jctr.assign((size_t)d->njobs * kJobCtr, 0);
    int reruns = 0;   // whole runs repeated with grown buffers
    for (int attempt = 0;; attempt++) {
        const uint32_t n_items = d->h_misc[M_NITEMS];
        const uint64_t nB = misc_nb(d);
        const uint64_t nres = d->h_ctr[C_RESULTS];
        bool again = false;
        if (n_items > d->cap_items) { d->cap_items = n_items + n_items / 4 + 1024; again = true; }
        // A larger boundary table only saves searches (the queries it could not plan
        // streamed or searched their ranges, so the run is complete): grow it within
        // the budget for the next run, and rerun now only for the first run of a
        // collection, whose other buffers are being sized anyway.
        const uint64_t want = std::min<uint64_t>(nB + nB / 4 + 1024, cfg.bounds);
        if (nB > d->cap_bounds && want > d->cap_bounds) {
            d->cap_bounds = (uint32_t)want;
            if (!d->sized) again = true;
        }
        if (d->h_ctr[C_OVERFLOW] > d->cap_ovf) {
            d->cap_ovf = (uint32_t)std::min<uint64_t>(d->h_ctr[C_OVERFLOW] * 5 / 4 + 1024, UINT32_MAX / 8);
            again = true;
        }
        if (nres > d->cap_out) {
            d->cap_out = (uint32_t)std::min<uint64_t>(nres + nres / 2 + 1024, UINT32_MAX / 16);
            again = true;
        }
        if (!again) break;
        if (attempt > 8) return fgss_fail(FGSS_ERR_DEVICE, "internal: buffer sizing did not converge");
        reruns++;
        int rc = launch_once(d, cfg, ops, use_graph);
        if (rc) return rc;
        HIPCK(hipStreamSynchronize(st));
    }
    d->sized = true;
    trace_mark("pipeline");
    T.graph = use_graph && ops.graph_ok;
    T.launches = 1 + reruns;
    T.index_entries = d->h_misc[M_NE];
    T.probe_tokens = d->h_misc[M_NP];
    {   // events recorded around the stages in both modes
        float ms = 0;
        if (hipEventElapsedTime(&ms, d->ev[0], d->ev[1]) == hipSuccess) T.ms_index = ms;
        if (hipEventElapsedTime(&ms, d->ev[1], d->ev[5]) == hipSuccess) T.ms_join = ms;
        if (hipEventElapsedTime(&ms, d->ev[4], d->ev[5]) == hipSuccess) T.ms_kernel = ms;
    }
    for (int k = C_ESCAN; k <= C_DPOST; k++) T.acc[k] += d->h_ctr[k];
    uint64_t nres = d->h_ctr[C_RESULTS], novf = d->h_ctr[C_OVERFLOW];
    tri.resize(3 * (size_t)nres);
    if (nres && nres <= d->out_copied) memcpy(tri.data(), d->h_out, (size_t)nres * 12);
    else if (nres) {
        HIPCK(hipMemcpyAsync(tri.data(), d->b_out.p, (size_t)nres * 12, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
    }
    d->out_hint = std::max<uint32_t>(d->out_hint, (uint32_t)std::min<uint64_t>(nres + nres / 4 + 1024, UINT32_MAX / 16));

    // overflowed work items: split each four ways and run them again
    const Item *cur = P<Item>(d->b_items);
    int round = 0;
    while (novf > 0) {
        if (++round > 48) return fgss_fail(FGSS_ERR_DEVICE, "internal: work-item splitting did not converge");
        const uint32_t ways = 4, cnt = (uint32_t)(novf * ways);
        int rc = ensure(d, d->b_items2, (size_t)cnt * sizeof(Item));
        if (rc) return rc;
        hipLaunchKernelGGL(k_resplit, dim3(grid_for((int64_t)novf, 256)), dim3(256), 0, st, cur, P<uint32_t>(d->b_ovf),
                           P<uint32_t>(d->b_ovfp),
                           (uint32_t)novf, ways, P<Item>(d->b_items2));
        HIPCK(hipGetLastError());
        std::swap(d->b_items, d->b_items2);   // the parent list stays alive in b_items2
        d->cap_gen++;
        cur = P<Item>(d->b_items);
        rc = ensure(d, d->b_ovf, (size_t)cnt * 4);
        if (rc) return rc;
        rc = ensure(d, d->b_ovfp, (size_t)cnt * 4);
        if (rc) return rc;
        // the round's launch adds to the per-job counters as its items complete: kept
        // before it, so that a launch repeated for a larger result buffer counts once
        const size_t jbytes = (size_t)d->njobs * kJobCtr * sizeof(uint64_t);
        rc = ensure(d, d->b_jsave, jbytes);
        if (rc) return rc;
        HIPCK(hipMemcpyAsync(d->b_jsave.p, d->b_jctr.p, jbytes, hipMemcpyDeviceToDevice, st));
        for (int tries = 0;; tries++) {
            if (tries > 8) return fgss_fail(FGSS_ERR_DEVICE, "internal: result buffer sizing did not converge");
            unsigned long long *ctr = P<unsigned long long>(d->b_ctr);
            uint32_t *misc = P<uint32_t>(d->b_misc);
            if (tries > 0) HIPCK(hipMemcpyAsync(d->b_jctr.p, d->b_jsave.p, jbytes, hipMemcpyDeviceToDevice, st));
            HIPCK(fill<unsigned long long>(ctr, C_COUNT, 0ull, st));
            HIPCK(fill<uint32_t>(misc + M_NRESPLIT, 1, cnt, st));
            JoinArgs A = ops.args();
            A.items = cur;
            A.n_items_dev = misc + M_NRESPLIT;
            A.n_items_cap = cnt;
            A.out = P<int32_t>(d->b_out);
            A.out_cap = d->cap_out;
            A.ovf = P<uint32_t>(d->b_ovf);
            A.ovfp = P<uint32_t>(d->b_ovfp);
            A.ovf_cap = cnt;   // a re-split item overflows at most once
            HIPCK(hipEventRecord(d->ev[4], st));
            rc = launch_join_cap(d, A, cfg.cap, false, ops.rs, false);   // re-split items are never dense
            if (rc) return rc;
            HIPCK(hipEventRecord(d->ev[5], st));
            HIPCK(hipMemcpyAsync(d->h_ctr, ctr, C_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
            HIPCK(hipStreamSynchronize(st));
            T.launches++;
            float ms = 0;
            HIPCK(hipEventElapsedTime(&ms, d->ev[4], d->ev[5]));
            T.ms_kernel += ms;
            if (d->h_ctr[C_RESULTS] <= d->cap_out) break;
            d->cap_out = (uint32_t)std::min<uint64_t>(d->h_ctr[C_RESULTS] * 2 + 1024, UINT32_MAX / 16);
            rc = ensure(d, d->b_out, (size_t)d->cap_out * 12);
            if (rc) return rc;
            d->cap_gen++;
        }
        for (int k = C_ESCAN; k <= C_DPOST; k++) T.acc[k] += d->h_ctr[k];
        nres = d->h_ctr[C_RESULTS];
        novf = d->h_ctr[C_OVERFLOW];
        const size_t old = tri.size();
        tri.resize(old + 3 * (size_t)nres);
        if (nres) {
            HIPCK(hipMemcpyAsync(tri.data() + old, d->b_out.p, (size_t)nres * 12, hipMemcpyDeviceToHost, st));
            HIPCK(hipStreamSynchronize(st));
        }
    }
    // per-job counters: the run's tail copied them unless rounds ran after it (the
    // items that overflowed had not flushed theirs)
    if (round > 0 || !d->jctr_async) {
        HIPCK(hipMemcpyAsync(d->h_jctr, d->b_jctr.p, (size_t)d->njobs * kJobCtr * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
    }
    for (size_t k = 0; k < T.jctr.size(); k++) T.jctr[k] = d->h_jctr[k];
    trace_mark("collect");
    return FGSS_OK;
}

// Run the join for the thresholds already in d->tj, synchronously.
static int run_pipeline(fgss_device_sets *d, const RunCfg &cfg, bool use_graph, std::vector<int32_t> &tri, RunTotals &T)
{
    tri.clear();
    T = RunTotals();
    T.jctr.assign((size_t)d->njobs * kJobCtr, 0);
    if (d->n == 0) return FGSS_OK;
    trace_mark("begin");
    const PipeOps ops = self_ops(d, cfg);
    int rc = launch_once(d, cfg, ops, use_graph);
    if (rc) return rc;
    trace_mark("enqueued");
    HIPCK(hipStreamSynchronize(d->st));
    trace_mark("synced");
    return finish_run(d, cfg, ops, use_graph, tri, T);
}

// Host threads for result post-processing: at most 16, the host share of one GPU
static unsigned host_workers()
{
    const unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1u : std::min(n, 16u);
}

// Order of the (q, s, overlap) triples by (q, s): an LSD radix sort (11-bit digits)
// of 64-bit keys (q, s) carried with the triple index; passes whose digit is
// constant are skipped.  out[k] low 32 bits = index of the k-th triple.
static void sort_pairs_by_key(const std::vector<int32_t> &tri, std::vector<uint64_t> &out)
{
    const size_t R = tri.size() / 3;
    // large result sets (batches) are sorted by up to 16 host threads: each counts and
    // scatters its own chunk, chunk after chunk per digit, so the sort stays stable
    const unsigned T = R >= ((size_t)1 << 17) ? host_workers() : 1u;
    // key = q above the bits of the largest s, so that the passes stop at the highest
    // bit any key holds
    uint32_t smax = 0, qmax = 0;
    for (size_t i = 0; i < R; i++) {
        qmax = std::max(qmax, (uint32_t)tri[3 * i]);
        smax = std::max(smax, (uint32_t)tri[3 * i + 1]);
    }
    int sbits = 0, kbits = 0;
    while (sbits < 32 && (smax >> sbits) != 0) sbits++;
    while (kbits < 32 && (qmax >> kbits) != 0) kbits++;
    kbits += sbits;
    std::vector<uint64_t> key(R), idx(R), k2(R), i2(R);
    uint64_t *kb[2] = {key.data(), k2.data()}, *ib[2] = {idx.data(), i2.data()};
    std::vector<std::array<size_t, 2048>> cnt(T);
    bool skip = false;   // written by thread 0 between barriers
    std::atomic<unsigned> arrived{0}, phase{0};
    auto barrier = [&](unsigned &mine) {
        const unsigned target = ++mine;
        if (arrived.fetch_add(1) + 1 == T * target) phase.store(target);
        while (phase.load() < target) std::this_thread::yield();
    };
    int final_buf = 0;
    auto worker = [&](unsigned t) {
        unsigned mine = 0;
        const size_t a = R * t / T, b = R * (t + 1) / T;
        for (size_t i = a; i < b; i++) {
            key[i] = ((uint64_t)(uint32_t)tri[3 * i] << sbits) | (uint32_t)tri[3 * i + 1];
            idx[i] = i;
        }
        int cur = 0;
        for (int shift = 0; shift < kbits; shift += 11) {
            const uint64_t *ks = kb[cur], *is = ib[cur];
            std::array<size_t, 2048> &c = cnt[t];
            c.fill(0);
            for (size_t i = a; i < b; i++) c[(ks[i] >> shift) & 2047]++;
            barrier(mine);
            if (t == 0) {   // starting position of (digit, chunk), digits in order, chunks in order
                size_t run = 0;
                skip = false;
                for (int dd = 0; dd < 2048; dd++) {
                    size_t tot = 0;
                    for (unsigned u = 0; u < T; u++) tot += cnt[u][dd];
                    if (tot == R) skip = true;   // one digit value only: the pass changes nothing
                    for (unsigned u = 0; u < T; u++) {
                        const size_t n_u = cnt[u][dd];
                        cnt[u][dd] = run;
                        run += n_u;
                    }
                }
            }
            barrier(mine);
            if (skip) continue;
            uint64_t *kd = kb[cur ^ 1], *id = ib[cur ^ 1];
            for (size_t i = a; i < b; i++) {
                const size_t at = c[(ks[i] >> shift) & 2047]++;
                kd[at] = ks[i];
                id[at] = is[i];
            }
            barrier(mine);
            cur ^= 1;
        }
        if (t == 0) final_buf = cur;
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; t++) th.emplace_back(worker, t);
    worker(0);
    for (auto &x : th) x.join();
    if (final_buf == 0) out.swap(idx);
    else out.swap(i2);
}

// Build a result object from triples (all with query index in [q0, q1) of the
// device record set, whose records q0.. are the records 0.. of collection c).
static int make_pairs(const fgss_collection *c, const std::vector<int32_t> &tri, const std::vector<uint64_t> &order,
                      size_t first, size_t count, int64_t q0, fgss_pairs **out)
{
    auto *p = (fgss_pairs *)calloc(1, sizeof(fgss_pairs));
    if (!p) return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    const size_t nb = std::max<size_t>(count, 1);
    p->q = (int32_t *)malloc(nb * 4);
    p->s = (int32_t *)malloc(nb * 4);
    p->rid = (int32_t *)malloc(nb * 4);
    p->sid = (int32_t *)malloc(nb * 4);
    p->overlap = (int32_t *)malloc(nb * 4);
    p->similarity = (float *)malloc(nb * 4);
    if (!p->q || !p->s || !p->rid || !p->sid || !p->overlap || !p->similarity) {
        fgss_pairs_free(p);
        return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    }
    for (size_t k = 0; k < count; k++) {
        const size_t i = (size_t)(order[first + k] & 0xFFFFFFFFull);
        const int32_t q = (int32_t)(tri[3 * i] - q0), s = (int32_t)(tri[3 * i + 1] - q0), ov = tri[3 * i + 2];
        p->q[k] = q;
        p->s[k] = s;
        p->rid[k] = c->id[(size_t)q];
        p->sid[k] = c->id[(size_t)s];
        p->overlap[k] = ov;
        p->similarity[k] = fgss_similarity((uint32_t)c->len[(size_t)q], (uint32_t)c->len[(size_t)s], (uint32_t)ov);
    }
    p->n_pairs = (int64_t)count;
    *out = p;
    return FGSS_OK;
}

static void fill_stats(const RunTotals &T, size_t job, fgss_join_stats &st)
{
    memset(&st, 0, sizeof(st));
    st.index_entries = T.index_entries;
    st.probe_tokens = T.probe_tokens;
    st.postings_scanned = (int64_t)T.jctr[job * kJobCtr + 0];
    st.candidates = (int64_t)T.jctr[job * kJobCtr + 1];
    st.intersections = (int64_t)T.jctr[job * kJobCtr + 2];
    st.results = (int64_t)T.jctr[job * kJobCtr + 3];
    st.verify_tokens = (int64_t)T.acc[C_VTOK];
    st.signature_checks = (int64_t)T.acc[C_SIGCHK];
    st.signature_checks_128 = (int64_t)T.acc[C_SIG2];
    st.verified = (int64_t)T.acc[C_VERIFIED];
    for (int k = 0; k < 6; k++) st.phase_cycles[k] = (int64_t)T.acc[C_CYC_SETUP + k];
    st.window_retries = (int64_t)T.acc[C_RETRY];
    st.dense_parts = (int64_t)T.acc[C_DENSE];
    st.dense_postings = (int64_t)T.acc[C_DPOST];
    st.ms_index = T.ms_index;
    st.ms_join = T.ms_join;
    st.ms_kernel = T.ms_kernel;
    st.kernel_launches = T.launches;
}

static int set_thresholds(fgss_device_sets *d, const float *t, uint32_t njobs)
{
    HIPCK(hipMemcpyAsync(d->tj, t, (size_t)njobs * sizeof(float), hipMemcpyHostToDevice, d->st));
    HIPCK(hipStreamSynchronize(d->st));
    return FGSS_OK;
}

// ---------------------------------------------------------------------------
// fgss_selfjoin
// ---------------------------------------------------------------------------
extern "C" int fgss_selfjoin(fgss_collection *c, float threshold, const fgss_join_opts *opts, fgss_pairs **out)
{
    if (!c || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(threshold > 0.0f) || !(threshold < 1.0f))   // fgssjoin_cli.cu:32-35
        return fgss_fail(FGSS_ERR_ARG, "Threshold must be a float in the range (0.0..1.0).");
    if (!c->jend.empty()) return fgss_fail(FGSS_ERR_ARG, "a batch collection is run with fgss_batch_run");
    const RunCfg cfg = make_cfg(opts);
    int rc = check_range(opts, (int64_t)c->len.size(), true);
    if (rc) return rc;
    rc = fgss_collection_upload(c);
    if (rc) return rc;
    auto t1 = std::chrono::steady_clock::now();
    fgss_device_sets *d = c->dev;
    CallScope scope(d, opts, true);
    // the one threshold travels as a kernel argument: no host copy, no synchronization
    HIPCK(fill<float>(d->tj, 1, threshold, d->st));
    std::vector<int32_t> tri;
    RunTotals T;
    // eager launches: replaying the stages as captured HIP graphs measured the same
    // (the dependent launches of a join cost a few microseconds each), and a one-off
    // join would pay for the capture
    rc = run_pipeline(d, cfg, false, tri, T);
    if (rc) return rc;
    auto t2 = std::chrono::steady_clock::now();
    std::vector<uint64_t> order;
    sort_pairs_by_key(tri, order);
    fgss_pairs *p = nullptr;
    rc = make_pairs(c, tri, order, 0, order.size(), 0, &p);
    if (rc) return rc;
    trace_mark("host_pairs");
    auto t3 = std::chrono::steady_clock::now();
    fill_stats(T, 0, p->stats);
    p->stats.ms_copy = std::chrono::duration<double, std::milli>(t3 - t2).count();
    p->stats.ms_total = std::chrono::duration<double, std::milli>(t3 - t1).count();
    *out = p;
    return FGSS_OK;
}

extern "C" void fgss_pairs_free(fgss_pairs *p)
{
    if (!p) return;
    free(p->q);
    free(p->s);
    free(p->rid);
    free(p->sid);
    free(p->overlap);
    free(p->similarity);
    free(p);
}

extern "C" int fgss_pairs_print(const fgss_pairs *p, void *file)
{
    if (!p || !file) return fgss_fail(FGSS_ERR_ARG, "null argument");
    FILE *f = (FILE *)file;
    for (int64_t k = 0; k < p->n_pairs; k++)
        fprintf(f, "%d %d %f\n", p->rid[k], p->sid[k], (double)p->similarity[k]);   // block.cu:193
    return FGSS_OK;
}

// Wait for all device work of this process (the reference synchronizes the whole
// device after every launch, e.g. src/fgssjoin/filtering.cu:38).
extern "C" int fgss_device_synchronize(void)
{
    HIPCK(hipDeviceSynchronize());
    return FGSS_OK;
}

// The self-join index at t, one line pair per term (print_inverted_index,
// src/fgssjoin/index.cu:218-243): "term: [t], size: n" then the ids of the records
// whose mid prefix holds the term, in internal order.
extern "C" int fgss_collection_print_index(fgss_collection *c, float threshold, void *file)
{
    if (!c || !file) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (!(threshold > 0.0f) || !(threshold < 1.0f))
        return fgss_fail(FGSS_ERR_ARG, "Threshold must be a float in the range (0.0..1.0).");
    if (!c->jend.empty()) return fgss_fail(FGSS_ERR_ARG, "a batch collection has one index per job");
    int rc = fgss_collection_upload(c);
    if (rc) return rc;
    fgss_device_sets *d = c->dev;
    rc = set_thresholds(d, &threshold, 1);
    if (!rc) rc = prepare_buffers(d);
    if (rc) return rc;
    uint32_t *misc = P<uint32_t>(d->b_misc);
    HIPCK(fill<uint32_t>(misc, M_COUNT, 0u, d->st));
    if (d->n > 0) {
        rc = enqueue_index(d, 0, 0u);
        if (rc) return rc;
    }
    HIPCK(hipMemcpyAsync(d->h_misc, misc, M_COUNT * 4, hipMemcpyDeviceToHost, d->st));
    HIPCK(hipStreamSynchronize(d->st));
    const uint32_t E = d->n > 0 ? d->h_misc[M_NE] : 0u, T = d->n_terms;
    std::vector<uint32_t> loff((size_t)T + 1, 0), psid((size_t)E);
    if (d->n > 0) {
        HIPCK(hipMemcpyAsync(loff.data(), d->b_loff.p, ((size_t)T + 1) * 4, hipMemcpyDeviceToHost, d->st));
        if (E) HIPCK(hipMemcpyAsync(psid.data(), d->b_psid.p, (size_t)E * 4, hipMemcpyDeviceToHost, d->st));
        HIPCK(hipStreamSynchronize(d->st));
    }
    FILE *f = (FILE *)file;
    for (uint32_t t = 0; t < T; t++) {
        const uint32_t a = loff[t], b = loff[t + 1];
        fprintf(f, "term: [%u], size: %u\n", t, b - a);
        for (uint32_t k = a; k < b; k++) fprintf(f, "(%u) ", (unsigned)c->id[psid[k]]);
        fprintf(f, "\n\n");
    }
    return FGSS_OK;
}

// ---------------------------------------------------------------------------
// batches: many independent self-joins (own collection, own threshold) run as
// one pipeline over the concatenated records -- one launch of every kernel for
// the whole batch, optionally replayed from a captured HIP graph
// ---------------------------------------------------------------------------
struct fgss_batch {
    std::vector<const fgss_collection *> colls;
    std::vector<float> t;
    std::vector<int64_t> rec_begin;   // first record of each job in the merged set
    fgss_collection merged;
    std::vector<int32_t> tri;         // last run: triples sorted by (q, s), merged indices
    std::vector<uint64_t> order;
    std::vector<size_t> job_first;    // last run: first sorted triple of each job
    // per job: its distinct record lengths and how many records have each, for the
    // job's own index-entry and probe-token counts (a run counts them for the batch)
    std::vector<std::vector<std::pair<uint32_t, int64_t>>> len_hist;
    RunTotals totals;
    bool has_run = false;
};

// Job j's "# Inverted Index Entries" and probe tokens at its threshold: the sums
// k_midlen's per-record prefix lengths make over the job's records.
static void batch_job_prefix_counts(const fgss_batch *b, size_t j, int64_t &entries, int64_t &probes)
{
    entries = probes = 0;
    const float t = b->t[j];
    for (const auto &lc : b->len_hist[j]) {
        const uint32_t l = lc.first;
        entries += (int64_t)std::min(fgss_mid_prefix(l, t), l) * lc.second;
        probes += (int64_t)std::min(fgss_max_prefix(l, t), l) * lc.second;
    }
}

static void batch_job_stats(const fgss_batch *b, size_t j, fgss_join_stats &st)
{
    fill_stats(b->totals, j, st);
    batch_job_prefix_counts(b, j, st.index_entries, st.probe_tokens);
}

extern "C" int fgss_batch_create(fgss_collection *const *colls, const float *thresholds, int32_t njobs,
                                 fgss_batch **out)
{
    if (!colls || !thresholds || !out || njobs < 1) return fgss_fail(FGSS_ERR_ARG, "null argument or no jobs");
    *out = nullptr;
    auto *b = new fgss_batch();
    fgss_collection &m = b->merged;
    int64_t nrec = 0, ntok = 0;
    for (int32_t j = 0; j < njobs; j++) {
        if (!colls[j]) { delete b; return fgss_fail(FGSS_ERR_ARG, "null collection for job %d", j); }
        if (!(thresholds[j] > 0.0f) || !(thresholds[j] < 1.0f)) {
            delete b;
            return fgss_fail(FGSS_ERR_ARG, "Threshold must be a float in the range (0.0..1.0).");
        }
        nrec += (int64_t)colls[j]->len.size();
        ntok += (int64_t)colls[j]->tok.size();
    }
    if (nrec >= (int64_t)INT32_MAX || ntok >= (int64_t)UINT32_MAX) {
        delete b;
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "batch exceeds 2^31 records or 2^32 tokens");
    }
    m.id.reserve((size_t)nrec);
    m.len.reserve((size_t)nrec);
    m.tok.reserve((size_t)ntok);
    m.job.reserve((size_t)nrec);
    int64_t term_off = 0;
    for (int32_t j = 0; j < njobs; j++) {   // every job's term range must fit int32 once offset
        term_off += colls[j]->n_terms;
        if (term_off >= (int64_t)INT32_MAX) {
            delete b;
            return fgss_fail(FGSS_ERR_UNSUPPORTED, "batch exceeds 2^31-1 distinct terms");
        }
    }
    term_off = 0;
    for (int32_t j = 0; j < njobs; j++) {
        const fgss_collection *c = colls[j];
        b->colls.push_back(c);
        b->t.push_back(thresholds[j]);
        b->rec_begin.push_back((int64_t)m.len.size());
        m.id.insert(m.id.end(), c->id.begin(), c->id.end());
        m.len.insert(m.len.end(), c->len.begin(), c->len.end());
        for (int32_t v : c->tok) m.tok.push_back((int32_t)(v + term_off));   // private term range per job
        m.job.insert(m.job.end(), c->len.size(), (uint32_t)j);
        m.jend.push_back((uint32_t)m.len.size());
        std::vector<int64_t> cnt((size_t)c->max_len + 1, 0);
        for (const int32_t l : c->len) cnt[(size_t)l]++;
        b->len_hist.emplace_back();
        for (size_t l = 0; l < cnt.size(); l++)
            if (cnt[l]) b->len_hist.back().emplace_back((uint32_t)l, cnt[l]);
        term_off += c->n_terms;
    }
    b->rec_begin.push_back((int64_t)m.len.size());
    (void)fgss_finish_collection(&m);   // the offset ids stay below 2^31-1 (checked above)
    m.n_terms = (int32_t)term_off;
    int rc = fgss_collection_upload(&m);
    if (rc) { fgss_device_release(&m); delete b; return rc; }
    *out = b;
    return FGSS_OK;
}

extern "C" void fgss_batch_free(fgss_batch *b)
{
    if (!b) return;
    fgss_device_release(&b->merged);
    delete b;
}

extern "C" int fgss_batch_set_thresholds(fgss_batch *b, const float *thresholds)
{
    if (!b || !thresholds) return fgss_fail(FGSS_ERR_ARG, "null argument");
    for (size_t j = 0; j < b->t.size(); j++) {
        if (!(thresholds[j] > 0.0f) || !(thresholds[j] < 1.0f))
            return fgss_fail(FGSS_ERR_ARG, "Threshold must be a float in the range (0.0..1.0).");
        b->t[j] = thresholds[j];
    }
    return FGSS_OK;
}

extern "C" int fgss_batch_run(fgss_batch *b, const fgss_join_opts *opts, int32_t use_graph, fgss_batch_stats *stats)
{
    if (!b) return fgss_fail(FGSS_ERR_ARG, "null batch");
    auto t0 = std::chrono::steady_clock::now();
    fgss_device_sets *d = b->merged.dev;
    const RunCfg cfg = make_cfg(opts);
    int rc = check_range(opts, d->n, false);
    if (rc) return rc;
    CallScope scope(d, opts, false);
    rc = set_thresholds(d, b->t.data(), (uint32_t)b->t.size());
    if (rc) return rc;
    const size_t nj = b->colls.size();
    const int groups = (int)std::min<size_t>((size_t)std::max(cfg.groups, 1), std::max<size_t>(nj, 1));
    if (groups <= 1) {
        rc = run_pipeline(d, cfg, use_graph != 0, b->tri, b->totals);
        if (rc) return rc;
    } else {
        // job groups of about equal record counts, each indexed and joined by a run of
        // its own over the same captured stages (the group's range is read on the device)
        b->tri.clear();
        b->totals = RunTotals();
        b->totals.jctr.assign(nj * kJobCtr, 0);
        std::vector<int32_t> tri;
        RunTotals T;
        const int64_t total = b->rec_begin[nj];
        size_t j0 = 0;
        for (int g = 0; g < groups && j0 < nj; g++) {
            size_t j1 = j0 + 1;
            const int64_t want = total * (g + 1) / groups;
            while (j1 < nj && (b->rec_begin[j1] < want || g == groups - 1)) j1++;
            d->grp_begin = (uint32_t)b->rec_begin[j0];
            d->grp_end = (uint32_t)b->rec_begin[j1];
            rc = run_pipeline(d, cfg, use_graph != 0, tri, T);
            d->grp_begin = d->grp_end = 0;
            if (rc) return rc;
            b->tri.insert(b->tri.end(), tri.begin(), tri.end());
            RunTotals &A = b->totals;
            for (int k = 0; k < C_COUNT; k++) A.acc[k] += T.acc[k];
            for (size_t k = 0; k < A.jctr.size(); k++) A.jctr[k] += T.jctr[k];
            A.index_entries += T.index_entries;
            A.probe_tokens += T.probe_tokens;
            A.ms_index += T.ms_index;
            A.ms_join += T.ms_join;
            A.ms_kernel += T.ms_kernel;
            A.launches += T.launches;
            A.graph = T.graph;
            j0 = j1;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    sort_pairs_by_key(b->tri, b->order);
    // jobs are contiguous record ranges, so the sorted triples are grouped by job
    b->job_first.assign(nj + 1, b->order.size());
    size_t k = 0;
    for (size_t j = 0; j < nj; j++) {
        while (k < b->order.size() && b->tri[3 * (size_t)(b->order[k] & 0xFFFFFFFFull)] < b->rec_begin[j]) k++;
        b->job_first[j] = k;
    }
    b->has_run = true;
    auto t2 = std::chrono::steady_clock::now();
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->ms_total = std::chrono::duration<double, std::milli>(t2 - t0).count();
        stats->ms_pipeline = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->n_pairs = (int64_t)b->order.size();
        for (size_t j = 0; j < nj; j++) stats->candidates += (int64_t)b->totals.jctr[j * kJobCtr + 1];
        stats->graph = b->totals.graph ? 1 : 0;
        stats->kernel_launches = b->totals.launches;
    }
    return FGSS_OK;
}

extern "C" int fgss_batch_pairs(fgss_batch *b, int32_t job, fgss_pairs **out)
{
    if (!b || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (!b->has_run) return fgss_fail(FGSS_ERR_ARG, "the batch has not been run");
    if (job < 0 || (size_t)job >= b->colls.size()) return fgss_fail(FGSS_ERR_ARG, "job %d out of range", job);
    const size_t first = b->job_first[(size_t)job], last = b->job_first[(size_t)job + 1];
    int rc = make_pairs(b->colls[(size_t)job], b->tri, b->order, first, last - first, b->rec_begin[(size_t)job], out);
    if (rc) return rc;
    batch_job_stats(b, (size_t)job, (*out)->stats);
    return FGSS_OK;
}

extern "C" int fgss_batch_pairs_all(fgss_batch *b, fgss_pairs **out, int64_t *job_first, fgss_join_stats *job_stats)
{
    if (!b || !out || !job_first) return fgss_fail(FGSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!b->has_run) return fgss_fail(FGSS_ERR_ARG, "the batch has not been run");
    const size_t nj = b->colls.size(), R = b->job_first[nj];
    auto *p = (fgss_pairs *)calloc(1, sizeof(fgss_pairs));
    if (!p) return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    const size_t nb = std::max<size_t>(R, 1);
    p->q = (int32_t *)malloc(nb * 4); p->s = (int32_t *)malloc(nb * 4); p->rid = (int32_t *)malloc(nb * 4);
    p->sid = (int32_t *)malloc(nb * 4); p->overlap = (int32_t *)malloc(nb * 4);
    p->similarity = (float *)malloc(nb * 4);
    if (!p->q || !p->s || !p->rid || !p->sid || !p->overlap || !p->similarity) {
        fgss_pairs_free(p);
        return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    }
    // jobs fill disjoint ranges: up to 16 host threads take them one at a time
    std::atomic<size_t> next{0};
    auto fill_jobs = [&] {
        for (size_t j; (j = next.fetch_add(1)) < nj;) {
            const fgss_collection *c = b->colls[j];
            const int64_t q0 = b->rec_begin[j];
            job_first[j] = (int64_t)b->job_first[j];
            for (size_t k = b->job_first[j]; k < b->job_first[j + 1]; k++) {
                const size_t i = (size_t)(b->order[k] & 0xFFFFFFFFull);
                const int32_t q = (int32_t)(b->tri[3 * i] - q0), sv = (int32_t)(b->tri[3 * i + 1] - q0);
                const int32_t ov = b->tri[3 * i + 2];
                p->q[k] = q;
                p->s[k] = sv;
                p->rid[k] = c->id[(size_t)q];
                p->sid[k] = c->id[(size_t)sv];
                p->overlap[k] = ov;
                p->similarity[k] =
                    fgss_similarity((uint32_t)c->len[(size_t)q], (uint32_t)c->len[(size_t)sv], (uint32_t)ov);
            }
            if (job_stats) {
                job_stats[j] = fgss_join_stats{};
                batch_job_stats(b, j, job_stats[j]);
            }
        }
    };
    const unsigned T = R >= ((size_t)1 << 17) ? std::min<unsigned>(host_workers(), (unsigned)nj) : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; t++) th.emplace_back(fill_jobs);
    fill_jobs();
    for (auto &x : th) x.join();
    job_first[nj] = (int64_t)R;
    p->n_pairs = (int64_t)R;
    *out = p;
    return FGSS_OK;
}

// Autotuner: time the batch under each (table slots, item load) choice, replaying
// captured graphs, and return the fastest.  reps runs per choice, the median counts.
// median of reps graph-replayed runs of the batch with options o (a first run captures
// the graphs and warms the buffers)
static int batch_time(fgss_batch *b, const fgss_join_opts &o, int32_t reps, double &ms)
{
    fgss_batch_stats st;
    int rc = fgss_batch_run(b, &o, 1, &st);
    if (rc) return rc;
    std::vector<double> v;
    for (int r = 0; r < reps; r++) {
        rc = fgss_batch_run(b, &o, 1, &st);
        if (rc) return rc;
        v.push_back(st.ms_total);
    }
    std::sort(v.begin(), v.end());
    ms = v[v.size() / 2];
    return FGSS_OK;
}

extern "C" int fgss_batch_autotune(fgss_batch *b, int32_t reps, fgss_join_opts *best, double *best_ms)
{
    if (!b || !best) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (reps < 1) reps = 3;
    double bms = 1e300;
    fgss_join_opts bo{};
    auto consider = [&](const fgss_join_opts &o) -> int {
        double ms = 0;
        const int rc = batch_time(b, o, reps, ms);
        if (rc) return rc;
        if (ms < bms) { bms = ms; bo = o; }
        return FGSS_OK;
    };
    // LDS table size x work-item load, then index groups and record blocks from the best
    for (int s : {512, 1024})
        for (int l : {50, 65, 80}) {
            fgss_join_opts o{};
            o.table_slots = s;
            o.load_pct = l;
            int rc = consider(o);
            if (rc) return rc;
        }
    const fgss_join_opts base = bo;
    for (int g : {2, 4, 8}) {
        if ((size_t)g > b->colls.size()) break;
        fgss_join_opts o = base;
        o.index_groups = g;
        int rc = consider(o);
        if (rc) return rc;
    }
    const fgss_join_opts base2 = bo;
    for (int r : {4096, 32768}) {
        fgss_join_opts o = base2;
        o.block_records = r;
        int rc = consider(o);
        if (rc) return rc;
    }
    *best = bo;
    if (best_ms) *best_ms = bms;
    return FGSS_OK;
}

// The tuning cache's key of a batch: its shape, thresholds, device and library version.
static std::string batch_tune_key(const fgss_batch *b)
{
    uint64_t h = 1469598103934665603ull;   // FNV-1a over the thresholds and job sizes
    auto mix = [&h](uint64_t v) {
        for (int k = 0; k < 8; k++) {
            h ^= (v >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    };
    int64_t toks = 0;
    for (size_t j = 0; j < b->colls.size(); j++) {
        uint32_t tb;
        memcpy(&tb, &b->t[j], 4);
        mix(tb);
        mix((uint64_t)b->colls[j]->len.size());
        toks += (int64_t)b->colls[j]->tok.size();
    }
    hipDeviceProp_t prop{};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) prop.gcnArchName[0] = 0;
    char buf[256];
    snprintf(buf, sizeof(buf), "jobs=%zu,records=%lld,tokens=%lld,h=%016llx,dev=%s/%d,lib=%s", b->colls.size(),
             (long long)b->rec_begin.back(), (long long)toks, (unsigned long long)h, prop.gcnArchName,
             prop.multiProcessorCount, fgss_version());
    for (char *c = buf; *c; c++)
        if (*c == ' ' || *c == '\t' || *c == '\n') *c = '_';
    return buf;
}

extern "C" int fgss_batch_autotune_cached(fgss_batch *b, int32_t reps, const char *cache_path, fgss_join_opts *best,
                                          double *best_ms, int32_t *from_cache)
{
    if (!b || !best || !cache_path) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (from_cache) *from_cache = 0;
    const std::string key = batch_tune_key(b);
    // a line: <key> table_slots load_pct index_groups block_records ms
    if (FILE *f = fopen(cache_path, "r")) {
        char line[512];
        bool hit = false;
        while (!hit && fgets(line, sizeof(line), f)) {
            char k[300];
            int s = 0, l = 0, g = 0, r = 0;
            double ms = 0;
            if (sscanf(line, "%299s %d %d %d %d %lf", k, &s, &l, &g, &r, &ms) == 6 && key == k &&
                (s == 256 || s == 512 || s == 1024 || s == 2048 || s == 4096) && l >= 0 && l <= 100 && g >= 0 &&
                r >= 0) {
                fgss_join_opts o{};
                o.table_slots = s;
                o.load_pct = l;
                o.index_groups = g;
                o.block_records = r;
                *best = o;
                if (best_ms) *best_ms = ms;
                hit = true;
            }
        }
        fclose(f);
        if (hit) {
            if (from_cache) *from_cache = 1;
            return FGSS_OK;
        }
    }
    double ms = 0;
    int rc = fgss_batch_autotune(b, reps, best, &ms);
    if (rc) return rc;
    if (best_ms) *best_ms = ms;
    FILE *f = fopen(cache_path, "a");
    if (!f) return fgss_fail(FGSS_ERR_IO, "cannot append to the tuning cache \"%s\"", cache_path);
    const bool ok = fprintf(f, "%s %d %d %d %d %.6f\n", key.c_str(), best->table_slots, best->load_pct,
                            best->index_groups, best->block_records, ms) > 0;
    if (fclose(f) != 0 || !ok) return fgss_fail(FGSS_ERR_IO, "cannot append to the tuning cache \"%s\"", cache_path);
    return FGSS_OK;
}

// ---------------------------------------------------------------------------
// serving: R-S joins of probe records against a resident indexed collection,
// with continuous batching.  No reference counterpart (the reference's
// "dualjoin" is only a usage string, src/preprocess_cli.cpp:26).
//
// R-S semantics: the self-join's closed form with the longer record of each pair as
// the query (the probe on ties).  Probe r and indexed record s form a result at
// threshold t iff the 4-way size filter holds, m = |R[0,qp) n S[0,sp)| > 0 with
// (qp, sp) = (max_prefix(|R|), mid_prefix(|S|)) when |S| <= |R| and
// (mid_prefix(|R|), max_prefix(|S|)) when |S| > |R|, the positional filter passes at
// the last prefix match, |R n S| >= floor(min_overlap) and the float32 Jaccard is
// >= t.  So an R-S join of a collection with itself has the self-join as its q < s
// half.  Each probe runs as two half-probes over the two parts of its size band.
// The index holds max_prefix(|S|, t_min) tokens per record, which covers every
// t >= t_min; postings past the source's own prefix at t are skipped.
// ---------------------------------------------------------------------------

// The largest partner length ls >= l with size_ok(l, ls) (at most max_len).
__device__ __forceinline__ uint32_t partner_lmax(uint32_t l, float t, uint32_t max_len)
{
    uint32_t top = fgss_floor_max_size(l, t);   // the filter needs ls <= floor(l / t)
    if (top > max_len) top = max_len;
    if (top < l) return l;
    uint32_t a = l + 1, b = top + 1;            // first failing length in [l + 1, top + 1)
    while (a < b) {
        uint32_t mid = (a + b) >> 1;
        if (fgss_size_ok(l, mid, t)) a = mid + 1;
        else b = mid;
    }
    return a - 1;
}

// Probe-prefix lengths of the two half-probes of every probe r: 2r probes the
// sources longer than r with r's mid prefix, 2r + 1 the others with its max prefix.
__global__ void k_probelen_rs(const uint32_t *__restrict__ len, const uint32_t *__restrict__ job,
                              const float *__restrict__ tj, uint32_t *__restrict__ probelen, int64_t np)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * np) return;
    const int64_t r = v >> 1;
    const uint32_t l = len[r];
    const float t = tj[job[r]];
    const uint32_t p = (v & 1) ? fgss_max_prefix(l, t) : fgss_mid_prefix(l, t);
    probelen[v] = p < l ? p : l;
}

// Per half-probe (kPlanGroup lanes each): its source window -- indexed records in the
// size-filter band of the probe, longer than it (2r) or not (2r + 1); a contiguous id
// range because lengths are non-increasing -- and each probe token's posting range
// inside it.
__global__ __launch_bounds__(256) void k_plan_rs(const uint32_t *__restrict__ ptok, const uint32_t *__restrict__ ppos,
                                                 const uint32_t *__restrict__ plen, int64_t np,
                                                 const uint32_t *__restrict__ pjob, const float *__restrict__ tj,
                                                 const uint32_t *__restrict__ probelen,
                                                 const uint32_t *__restrict__ poff, const uint32_t *__restrict__ slen,
                                                 uint32_t ns, uint32_t smax_len, const uint32_t *__restrict__ loff,
                                                 const uint32_t *__restrict__ psid, const uint32_t *__restrict__ fence,
                                                 uint32_t n_terms, uint32_t target, uint2 *__restrict__ rng,
                                                 uint32_t *__restrict__ wlo_out, uint32_t *__restrict__ hi_out,
                                                 uint32_t *__restrict__ nparts, uint32_t *__restrict__ eq_out,
                                                 uint64_t *__restrict__ bcnt, uint32_t stream_parts)
{
    const int64_t ngroups = (int64_t)gridDim.x * (blockDim.x / kPlanGroup);
    const uint32_t gl = (uint32_t)lane_id() & (kPlanGroup - 1);   // lane within the group
    // whole wavefronts iterate together (see k_plan)
    const int64_t nv = 2 * np, nround = (nv + kWave / kPlanGroup - 1) / (kWave / kPlanGroup) * (kWave / kPlanGroup);
    for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kPlanGroup; g < nround; g += ngroups) {
        const bool live = g < nv;
        const int64_t v = live ? g : nv - 1;
        const int64_t r = v >> 1;
        const bool longer = (v & 1) == 0;
        const uint32_t lr = plen[r], qp = live ? probelen[v] : 0u;
        const float t = tj[pjob[r]];
        // band [lmin, lmax] split at lr: (lr, lmax] for 2r, [lmin, lr] for 2r + 1
        const uint32_t top = longer ? partner_lmax(lr, t, smax_len) : lr;
        const uint32_t bot = longer ? lr + 1 : partner_lmin(lr, t);
        uint32_t a = 0, b = ns;   // first source with length <= top
        while (a < b) {
            uint32_t mid = (a + b) >> 1;
            if (slen[mid] > top) a = mid + 1;
            else b = mid;
        }
        const uint32_t wlo = a;
        b = ns;                   // first source with length < bot
        while (a < b) {
            uint32_t mid = (a + b) >> 1;
            if (slen[mid] >= bot) a = mid + 1;
            else b = mid;
        }
        const uint32_t whi = a;
        const uint32_t ro = poff[v];
        uint32_t e = 0;
        for (uint32_t j = gl; j < qp; j += kPlanGroup) {
            const uint32_t term = ptok[ppos[r] + j];
            uint32_t x = 0, y = 0;
            if (term < n_terms && wlo < whi) {   // a q-gram the index has never seen has no postings
                const uint32_t L0 = loff[term], L1 = loff[term + 1];
                lower_bound_sid2(psid, fence, L0, L1, wlo, L0, L1, whi, x, y);   // wlo < whi here
            }
            rng[ro + j] = make_uint2(x, y);
            e += y - x;
        }
#pragma unroll
        for (int d = kPlanGroup / 2; d > 0; d >>= 1) e += __shfl_xor(e, d, kPlanGroup);
        if (live && gl == 0) {
            wlo_out[v] = wlo;
            hi_out[v] = whi;
            eq_out[v] = e;
            const uint32_t npt = e == 0 ? 0u : (e + target - 1) / target;
            nparts[v] = npt;
            bcnt[v] = npt > 1 && (npt <= stream_parts || qp > kStreamQp) ? (uint64_t)(npt - 1) * qp : 0ull;
        }
    }
}

// host-side request and server state
constexpr uint32_t kReserveJobs = 4096;   // requests per run a slot is sized for at creation
constexpr int64_t kServeBudget = (int64_t)256 << 20;   // default budget: device bytes of the runs in flight

// Device memory a request occupies in a run's work buffers, charged against the server's
// budget: its staged records (4 B a token, 12 B a probe), the plan arrays of its two
// half-probes (48 B each), the posting ranges of their prefix tokens (8 B each), and its
// estimated work items (48 B each) with their boundary rows (4 B per item and probe-prefix
// token of an average probe).  The items are estimated from the postings the probe
// prefixes may reach (prefix tokens x the index's mean list length) in parts of `target`
// postings, plus one per half-probe.  The candidates themselves live in LDS: the work
// items and boundary rows are the buffers that grow with them.
static int64_t serve_request_bytes(const fgss_collection *probes, float t, double mean_list, uint32_t target)
{
    double tok = 0, ptok = 0;
    for (int32_t l : probes->len) {
        tok += (double)l;
        ptok += (double)std::min<uint32_t>(fgss_max_prefix((uint32_t)l, t), (uint32_t)l);
    }
    const double np = (double)probes->len.size();
    const double items = ptok * mean_list / std::max<uint32_t>(target, 1u) + 2.0 * np;
    const double qp = np > 0 ? ptok / np : 0.0;
    return (int64_t)(4.0 * tok + 12.0 * np + 96.0 * np + 16.0 * ptok + 48.0 * items + 4.0 * items * qp) + 4096;
}

struct ServeRequest {
    int64_t id = 0;
    const fgss_collection *probes = nullptr;
    float t = 0.f;
    int64_t est = 0;                    // device bytes charged against the budget (serve_request_bytes)
    bool done = false;
    int status = FGSS_OK;
    std::vector<int32_t> tri;           // (probe, source, overlap), probe indices local to the request
    uint64_t cnt[kJobCtr] = {0, 0, 0, 0};
    int32_t launches = 0;               // join-kernel launches of its run
    int32_t batch_requests = 0;         // requests that shared its run
    int64_t inflight = 0;               // charges in flight once its run was admitted
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
};

struct ServeSlot {
    fgss_device_sets *d = nullptr;      // probe records + workspace (own stream)
    DevBuf b_in;                        // the run's probes: StageLayout arrays
    DevBuf b_sig, b_sig64, b_sigbad, b_wlo;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<ServeRequest *> reqs;
    std::vector<int64_t> first;         // first probe of each request in the merged set
    uint32_t *h_in = nullptr;           // pinned staging of the merged probes (StageLayout)
    size_t h_in_cap = 0;                // its words
    size_t jctr_cap = 0;                // jobs the pinned counter mirror holds
    int64_t est = 0;
    uint64_t seq = 0;                   // launch order
};

struct fgss_server {
    fgss_device_sets *idx = nullptr;    // private device copy of the indexed records + the R-S index
    const fgss_collection *coll = nullptr;
    float t_min = 0.f;
    int64_t budget = 0, inflight = 0;
    double mean_list = 1.0;             // index entries per term (for the estimate)
    RunCfg cfg;
    std::vector<ServeSlot> slots;
    std::deque<ServeRequest *> pending;
    std::map<int64_t, ServeRequest *> reqs;
    int64_t next_id = 1;
    int64_t max_probes = 1 << 16;       // probes per slot run
    uint64_t launch_seq = 0;
    uint32_t index_entries = 0;
    bool hold = false;                  // admission paused
    bool empty = false;                 // no indexed tokens: every request has no pairs
};

static int slot_prepare(ServeSlot &sl)
{
    fgss_device_sets *d = sl.d;
    // two half-probes per probe; their prefixes hold at most 2 x the probe tokens
    const int64_t v = 2 * std::max<int64_t>(d->n, 1), E = 2 * std::max<int64_t>(d->ntok, 1);
    // capacities follow the run's size (items >= half-probes); data-dependent growth
    // beyond that is found after the run, which then runs again
    d->cap_items = std::max<uint32_t>(d->cap_items, (uint32_t)std::min<int64_t>(v + 4096, UINT32_MAX / 2));
    d->cap_bounds = std::max<uint32_t>(d->cap_bounds, (uint32_t)std::min<int64_t>(E + 4096, kBoundsBudget));
    if (d->cap_out == 0) d->cap_out = 65536;
    GROW(d->b_probelen, (size_t)v * 4);
    GROW(d->b_poff, (size_t)(v + 1) * 4);
    GROW(d->b_scanscr, (size_t)scan_scratch_elems(v + 1) * 8);   // also 64-bit scans
    GROW(d->b_rng, (size_t)E * sizeof(uint2));
    GROW(sl.b_wlo, (size_t)v * 4);
    GROW(d->b_hi, (size_t)v * 4);
    GROW(d->b_eq, (size_t)v * 4);
    GROW(d->b_nparts, (size_t)v * 4);
    GROW(d->b_partoff, (size_t)(v + 1) * 4);
    GROW(d->b_bcnt, (size_t)v * 8);
    GROW(d->b_bofs, (size_t)(v + 1) * 8);
    GROW(d->b_nitems, (size_t)v * 4);
    d->cap_ovf = std::max(d->cap_ovf, d->cap_items);
    GROW(d->b_ovfp, (size_t)d->cap_ovf * 4);
    {
        int rc_ = wave_slots(d);
        if (rc_) return rc_;
    }
    GROW(d->b_bounds, (size_t)d->cap_bounds * 4);
    GROW(d->b_items, (size_t)d->cap_items * sizeof(Item));
    GROW(d->b_ovf, (size_t)d->cap_ovf * 4);
    GROW(d->b_out, (size_t)d->cap_out * 12);
    GROW(d->b_ctr, C_COUNT * sizeof(uint64_t));
    GROW(d->b_jctr, (size_t)std::max<uint32_t>(d->njobs, 1) * kJobCtr * sizeof(uint64_t));
    GROW(d->b_misc, M_COUNT * 4);
    return FGSS_OK;
}

static JoinArgs rs_args(fgss_server *srv, ServeSlot &sl)
{
    fgss_device_sets *d = sl.d, *x = srv->idx;
    JoinArgs A = join_args(d);
    A.tok = x->tok; A.pos = x->pos; A.len = x->len;
    A.sig = x->sig; A.sig64 = x->sig64; A.sigbad = x->sigbad;
    A.loff = P<uint32_t>(x->b_loff); A.post = P<uint2>(x->b_post);
    A.psid = P<uint32_t>(x->b_psid); A.fence = P<uint32_t>(x->b_fence);
    A.qtok = d->tok; A.qsig = d->sig; A.qsig64 = d->sig64; A.qsigbad = d->sigbad;
    return A;
}

// Enqueue an R-S run of the slot's probes on the slot's stream (no host syncs).
static int enqueue_rs(fgss_server *srv, ServeSlot &sl, bool timing)
{
    fgss_device_sets *d = sl.d, *x = srv->idx;
    hipStream_t st = d->st;
    const int64_t n = d->n;
    const RunCfg &cfg = srv->cfg;
    uint32_t *misc = P<uint32_t>(d->b_misc), *scr = P<uint32_t>(d->b_scanscr);
    uint32_t *probelen = P<uint32_t>(d->b_probelen), *poff = P<uint32_t>(d->b_poff);
    uint32_t *hi = P<uint32_t>(d->b_hi), *eq = P<uint32_t>(d->b_eq), *nparts = P<uint32_t>(d->b_nparts),
             *partoff = P<uint32_t>(d->b_partoff), *wlo = P<uint32_t>(sl.b_wlo);
    uint64_t *bcnt = P<uint64_t>(d->b_bcnt), *bofs = P<uint64_t>(d->b_bofs), *scr64 = P<uint64_t>(d->b_scanscr);
    if (timing) HIPCK(hipEventRecord(d->ev[0], st));
    HIPCK(clear_run(d, st));
    const int64_t v = 2 * n;   // half-probes
    hipLaunchKernelGGL(k_probelen_rs, dim3(grid_for(v, 256)), dim3(256), 0, st, d->len, d->job, d->tj, probelen, n);
    HIPCK(hipGetLastError());
    HIPCK(exclusive_scan<uint32_t>(probelen, poff, v, misc + M_NP, scr, st));
    if (timing) HIPCK(hipEventRecord(d->ev[1], st));
    hipLaunchKernelGGL(k_plan_rs, dim3((unsigned)std::min<int64_t>(grid_for(v * kPlanGroup, 256), 4096)), dim3(256), 0, st,
                       d->tok, d->pos, d->len, n, d->job, d->tj, probelen, poff, x->len, (uint32_t)x->n,
                       (uint32_t)x->max_len, P<uint32_t>(x->b_loff), P<uint32_t>(x->b_psid), P<uint32_t>(x->b_fence),
                       x->n_terms, cfg.target, P<uint2>(d->b_rng), wlo, hi, nparts, eq, bcnt, cfg.stream_parts);
    HIPCK(hipGetLastError());
    const uint32_t blim = std::min(d->cap_bounds, cfg.bounds);
    uint32_t *nitems = P<uint32_t>(d->b_nitems);
    HIPCK(exclusive_scan<uint64_t>(bcnt, bofs, v, (uint64_t *)(misc + M_NB64), scr64, st));
    hipLaunchKernelGGL(k_nitems, dim3(grid_for(v, 256)), dim3(256), 0, st, nparts, probelen, bofs, bcnt, blim, v, nitems);
    HIPCK(hipGetLastError());
    HIPCK(exclusive_scan<uint32_t>(nitems, partoff, v, misc + M_NITEMS, scr, st));
    hipLaunchKernelGGL(k_bounds, dim3(kBoundsGrid), dim3(256), 0, st, v, probelen, poff, P<uint2>(d->b_rng), wlo, hi,
                       nparts, bofs, bcnt, (const uint64_t *)(misc + M_NB64), P<uint32_t>(x->b_psid),
                       P<uint32_t>(x->b_fence), P<uint32_t>(d->b_bounds), blim);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_items, dim3(kItemsGrid), dim3(256), 0, st, 1, wlo, hi, nparts, probelen, eq, partoff,
                       bofs, bcnt, blim, d->job, d->tj, d->len, d->pos, poff, v, misc + M_NITEMS, P<Item>(d->b_items),
                       d->cap_items, (const uint32_t *)nullptr);
    HIPCK(hipGetLastError());
    if (timing) HIPCK(hipEventRecord(d->ev[4], st));
    int rc = launch_join_cap(d, rs_args(srv, sl), cfg.cap, false, true);
    if (rc) return rc;
    if (timing) HIPCK(hipEventRecord(d->ev[5], st));
    rc = enqueue_copies(d);
    if (rc) return rc;
    HIPCK(hipEventRecord(sl.done, st));
    return FGSS_OK;
}

static PipeOps rs_ops(fgss_server *srv, ServeSlot &sl)
{
    PipeOps o;
    o.prepare = [&sl] { return slot_prepare(sl); };
    o.enqueue = [srv, &sl](bool timing) { return enqueue_rs(srv, sl, timing); };
    o.args = [srv, &sl] { return rs_args(srv, sl); };
    o.rs = true;
    o.graph_ok = false;
    return o;
}

// The merged probes of a slot run as one block of 32-bit words, staged in pinned host
// memory and copied to the device in one asynchronous copy: the tokens, then per
// probe its length, token offset and request, then per request its end and
// threshold, each array starting on a 16-byte boundary.
struct StageLayout {
    size_t tok = 0, len = 0, pos = 0, job = 0, jend = 0, tj = 0, words = 0;
    StageLayout(int64_t n, int64_t ntok, uint32_t jobs)
    {
        auto up = [](size_t x) { return (x + 3) & ~(size_t)3; };
        len = up((size_t)ntok);
        pos = len + up((size_t)n);
        job = pos + up((size_t)n);
        jend = job + up((size_t)n);
        tj = jend + up(jobs);
        words = std::max<size_t>(tj + up(jobs), 4);
    }
};

// Grow a slot's pinned staging (doubling; sized at server creation, so that runs
// within the reservation never free pinned memory, which waits for the device).  The
// slot reuses it only after the run that copied from it has retired.
static int stage_reserve(ServeSlot &sl, size_t words)
{
    if (sl.h_in_cap >= words) return FGSS_OK;
    retire_host(sl.d, sl.h_in);
    sl.h_in = nullptr;
    const size_t cap = std::max<size_t>({words, 2 * sl.h_in_cap, (size_t)1 << 16});
    sl.h_in_cap = 0;
    HIPCK(hipHostMalloc((void **)&sl.h_in, cap * 4, hipHostMallocDefault));
    sl.h_in_cap = cap;
    return FGSS_OK;
}

// Copy the staged probes of a slot to the device (growing its arrays) and compute
// their signatures, asynchronously on the slot's stream.
static int slot_load(ServeSlot &sl, int64_t n, int64_t ntok, uint32_t nj, int32_t max_len, uint32_t n_terms)
{
    fgss_device_sets *d = sl.d;
    const StageLayout L(n, ntok, nj);
    GROW(sl.b_in, L.words * 4);
    GROW(sl.b_sig, (size_t)std::max<int64_t>(n, 1) * 4 * kSigWords);
    GROW(sl.b_sig64, (size_t)std::max<int64_t>(n, 1) * 4 * kSig64Words);
    GROW(sl.b_sigbad, (size_t)std::max<int64_t>(n, 1));
    uint32_t *in = P<uint32_t>(sl.b_in);
    d->tok = in + L.tok; d->len = in + L.len; d->pos = in + L.pos; d->job = in + L.job; d->jend = in + L.jend;
    d->tj = (float *)(in + L.tj);
    d->sig = P<uint32_t>(sl.b_sig); d->sig64 = P<uint32_t>(sl.b_sig64); d->sigbad = P<uint8_t>(sl.b_sigbad);
    d->n = n; d->ntok = ntok; d->njobs = nj; d->max_len = max_len; d->n_terms = n_terms;
    // pinned per-job counter mirror, grown to this batch's jobs
    if (sl.jctr_cap < (size_t)std::max<uint32_t>(nj, 1)) {
        retire_host(d, d->h_jctr);
        d->h_jctr = nullptr;
        sl.jctr_cap = std::max<size_t>({(size_t)nj, 64, sl.jctr_cap * 2});
        HIPCK(hipHostMalloc((void **)&d->h_jctr, sl.jctr_cap * kJobCtr * 8, hipHostMallocDefault));
    }
    hipStream_t st = d->st;
    HIPCK(hipMemcpyAsync(in, sl.h_in, L.words * 4, hipMemcpyHostToDevice, st));
    if (n) {
        const int64_t blocks = std::min<int64_t>((n + 3) / 4, 65536);
        hipLaunchKernelGGL(k_signature, dim3((unsigned)blocks), dim3(256), 0, st, d->tok, d->pos, d->len, n, d->sig,
                           d->sig64, d->sigbad);
        HIPCK(hipGetLastError());
    }
    return FGSS_OK;
}

// Size a slot, when the server is created, for runs of up to n probes holding ntok
// tokens and `jobs` requests, so that such runs never reallocate: a reallocation
// waits for the whole device (stalling the runs of the other slots) and the run that
// found a buffer too small is enqueued again.
static int slot_reserve(ServeSlot &sl, int64_t n, int64_t ntok, uint32_t jobs, int64_t items)
{
    fgss_device_sets *d = sl.d;
    d->cap_items = std::max<uint32_t>(d->cap_items, (uint32_t)std::min<int64_t>(items, UINT32_MAX / 2));
    const StageLayout L(n, ntok, jobs);
    GROW(sl.b_in, L.words * 4);
    {
        const int rc_ = stage_reserve(sl, L.words);
        if (rc_) return rc_;
    }
    GROW(sl.b_sig, (size_t)n * 4 * kSigWords);
    GROW(sl.b_sig64, (size_t)n * 4 * kSig64Words);
    GROW(sl.b_sigbad, (size_t)n);
    if (sl.jctr_cap < jobs) {
        retire_host(d, d->h_jctr);
        d->h_jctr = nullptr;
        sl.jctr_cap = jobs;
        HIPCK(hipHostMalloc((void **)&d->h_jctr, sl.jctr_cap * kJobCtr * 8, hipHostMallocDefault));
    }
    if (d->h_out_cap < (1u << 18)) {   // pinned mirror of a run's first 2^18 result triples (3 MB)
        retire_host(d, d->h_out);
        d->h_out = nullptr;
        d->h_out_cap = 0;
        HIPCK(hipHostMalloc((void **)&d->h_out, (size_t)(1u << 18) * 12, hipHostMallocDefault));
        d->h_out_cap = 1u << 18;
    }
    d->cap_out = std::max<uint32_t>(d->cap_out, 1u << 20);
    // boundary rows grow with the probes' posting reach, not their tokens: 2^26 entries
    // (256 MB) plan the heaviest runs of a 2M-title index without growing
    d->cap_bounds = std::max<uint32_t>(d->cap_bounds, std::min<uint32_t>(1u << 26, kBoundsBudget));
    d->sized = true;   // runs that outgrow a reserved capacity still complete
    // slot_prepare sizes the run buffers for d's record counts: the reserved ones for now
    const int64_t n0 = d->n, ntok0 = d->ntok;
    const uint32_t nj0 = d->njobs;
    d->n = n; d->ntok = ntok; d->njobs = jobs;
    const int rc = slot_prepare(sl);
    d->n = n0; d->ntok = ntok0; d->njobs = nj0;
    return rc;
}

// Merge the slot's requests into one probe set (request k = job k, its own
// threshold) and enqueue its run on the slot's stream.
static int slot_launch(fgss_server *srv, ServeSlot &sl)
{
    trace_mark("begin");
    int64_t n = 0, ntok = 0;
    for (const ServeRequest *r : sl.reqs) {
        n += (int64_t)r->probes->len.size();
        ntok += (int64_t)r->probes->tok.size();
    }
    const uint32_t nj = (uint32_t)sl.reqs.size();
    const StageLayout L(n, ntok, nj);
    {
        const int rc_ = stage_reserve(sl, L.words);
        if (rc_) return rc_;
    }
    uint32_t *tok = sl.h_in + L.tok, *len = sl.h_in + L.len, *pos = sl.h_in + L.pos, *job = sl.h_in + L.job,
             *jend = sl.h_in + L.jend;
    float *tj = (float *)(sl.h_in + L.tj);
    sl.first.clear();
    const auto now = std::chrono::steady_clock::now();
    int64_t i = 0;
    uint32_t p = 0;   // token offset of probe i (records are concatenated in order)
    int32_t mx = 0;
    for (uint32_t k = 0; k < nj; k++) {
        const fgss_collection *c = sl.reqs[k]->probes;
        sl.first.push_back(i);
        if (!c->tok.empty()) memcpy(tok + p, c->tok.data(), c->tok.size() * 4);
        for (const int32_t l : c->len) {
            len[i] = (uint32_t)l;
            pos[i] = p;
            job[i] = k;
            p += (uint32_t)l;
            mx = std::max(mx, l);
            i++;
        }
        jend[k] = (uint32_t)i;
        tj[k] = sl.reqs[k]->t;
        sl.reqs[k]->t_start = now;
        sl.reqs[k]->batch_requests = (int32_t)nj;
    }
    sl.first.push_back(i);
    trace_mark("srv_merge");
    int rc = slot_load(sl, n, ntok, nj, mx, (uint32_t)srv->idx->n_terms);
    if (rc) return rc;
    trace_mark("srv_load");
    if (sl.d->n > 0 && !srv->empty) {
        rc = slot_prepare(sl);
        if (rc) return rc;
        rc = enqueue_rs(srv, sl, true);
        if (rc) return rc;
        trace_mark("srv_enqueue");
    } else {
        HIPCK(hipEventRecord(sl.done, sl.d->st));   // nothing to join
    }
    sl.busy = true;
    return FGSS_OK;
}

// A slot's run has completed on the device: finish it on the host (reruns with
// grown capacities, rounds for overflowed items), then hand every request its
// (probe, source, overlap) triples and counters.
static int slot_retire(fgss_server *srv, ServeSlot &sl)
{
    std::vector<int32_t> tri;
    RunTotals T;
    int rc = FGSS_OK;
    trace_mark("srv_ready");
    if (sl.d->n > 0 && !srv->empty) {
        const PipeOps ops = rs_ops(srv, sl);
        rc = finish_run(sl.d, srv->cfg, ops, false, tri, T);
        trace_mark("srv_finish");
    } else {
        T.jctr.assign(sl.reqs.size() * kJobCtr, 0);
    }
    const auto now = std::chrono::steady_clock::now();
    std::vector<std::vector<int32_t>> per(sl.reqs.size());
    if (rc == FGSS_OK) {
        for (size_t i = 0; i < tri.size() / 3; i++) {
            const int32_t q = tri[3 * i];
            const size_t k =
                (size_t)(std::upper_bound(sl.first.begin(), sl.first.end(), (int64_t)q) - sl.first.begin()) - 1;
            per[k].push_back(q - (int32_t)sl.first[k]);
            per[k].push_back(tri[3 * i + 1]);
            per[k].push_back(tri[3 * i + 2]);
        }
    }
    for (size_t k = 0; k < sl.reqs.size(); k++) {
        ServeRequest *r = sl.reqs[k];
        r->status = rc;
        r->tri.swap(per[k]);
        for (int c = 0; c < kJobCtr; c++) r->cnt[c] = rc == FGSS_OK ? T.jctr[k * kJobCtr + c] : 0;
        r->launches = T.launches;
        r->done = true;
        r->t_done = now;
    }
    srv->inflight -= sl.est;
    sl.est = 0;
    sl.reqs.clear();
    sl.busy = false;
    trace_mark("srv_split");
    return rc;
}

// Retire completed slots, then admit pending requests into idle slots, first come
// first served, as many per slot as the work budget and the per-slot
// probe limit allow (at least one when nothing is in flight, so every request is
// eventually admitted).  With block set, waits for the oldest busy slot first.
static int server_progress(fgss_server *srv, bool block)
{
    if (block) {
        ServeSlot *oldest = nullptr;
        for (ServeSlot &sl : srv->slots)
            if (sl.busy && (!oldest || sl.seq < oldest->seq)) oldest = &sl;
        if (oldest) HIPCK(hipEventSynchronize(oldest->done));
    }
    for (ServeSlot &sl : srv->slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return fgss_fail(FGSS_ERR_DEVICE, "slot event: %s", hipGetErrorString(q));
        slot_retire(srv, sl);   // a failed run is reported through its requests
    }
    for (ServeSlot &sl : srv->slots) {
        if (srv->hold || srv->pending.empty()) break;
        if (sl.busy) continue;
        int64_t probes = 0;
        while (!srv->pending.empty()) {
            ServeRequest *r = srv->pending.front();
            const int64_t np = (int64_t)r->probes->len.size();
            const bool fits = srv->inflight + r->est <= srv->budget && probes + np <= srv->max_probes;
            const bool alone = sl.reqs.empty() && srv->inflight == 0;
            if (!fits && !alone) break;
            srv->pending.pop_front();
            sl.reqs.push_back(r);
            sl.est += r->est;
            srv->inflight += r->est;
            probes += np;
            if (!fits) break;
        }
        if (sl.reqs.empty()) break;
        for (ServeRequest *r : sl.reqs) r->inflight = srv->inflight;
        sl.seq = srv->launch_seq++;
        int rc = slot_launch(srv, sl);
        if (rc) {
            for (ServeRequest *r : sl.reqs) { r->status = rc; r->done = true; r->t_done = std::chrono::steady_clock::now(); }
            srv->inflight -= sl.est;
            sl.est = 0;
            sl.reqs.clear();
            (void)hipStreamSynchronize(sl.d->st);
            return rc;
        }
    }
    return FGSS_OK;
}

extern "C" int fgss_server_create(const fgss_collection *indexed, float t_min, int32_t slots, int64_t budget_bytes,
                                  const fgss_join_opts *opts, fgss_server **out)
{
    if (!indexed || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(t_min > 0.0f) || !(t_min < 1.0f)) return fgss_fail(FGSS_ERR_ARG, "t_min must be a float in (0.0..1.0)");
    if (slots < 1 || slots > 16) return fgss_fail(FGSS_ERR_ARG, "slots must be in 1..16");
    if (opts && opts->table_slots && opts->table_slots != 256 && opts->table_slots != 512 &&
        opts->table_slots != 1024 && opts->table_slots != 2048 && opts->table_slots != 4096)
        return fgss_fail(FGSS_ERR_ARG, "table_slots must be 256, 512, 1024, 2048 or 4096");
    if (!indexed->jend.empty()) return fgss_fail(FGSS_ERR_ARG, "the indexed collection must be a plain collection");
    // its size bands are contiguous id ranges: lengths must be non-increasing
    if (!indexed->sorted_desc)
        return fgss_fail(FGSS_ERR_UNSUPPORTED,
                         "the indexed collection's lengths must be non-increasing in internal order (a size-ascending .bin)");
    int rc0 = check_range(opts, (int64_t)indexed->len.size(), false);
    if (rc0) return rc0;
    auto *srv = new fgss_server();
    srv->coll = indexed;
    srv->t_min = t_min;
    srv->budget = budget_bytes > 0 ? budget_bytes : kServeBudget;
    srv->cfg = make_cfg(opts);
    int rc = upload_records(&srv->idx, indexed);
    if (rc) { delete srv; return rc; }
    fgss_device_sets *x = srv->idx;
    srv->empty = x->n == 0 || x->ntok == 0;
    // the R-S index: the first max_prefix(|s|, t_min) tokens of every record (built on
    // the caller's stream when opts names one; the slots' runs use streams of their own,
    // so that they overlap)
    {
        CallScope scope(x, opts, false);   // restored before x can be released below
        rc = set_thresholds(x, &t_min, 1);
        if (!rc) rc = prepare_buffers(x);
        if (!rc && srv->empty) x->h_misc[M_NE] = 0;
        if (!rc && !srv->empty) {
            uint32_t *misc = P<uint32_t>(x->b_misc);
            if (fill<uint32_t>(misc, M_COUNT, 0u, x->st) != hipSuccess) rc = fgss_fail(FGSS_ERR_DEVICE, "fill failed");
            if (!rc) rc = enqueue_index(x, 1, 0u);
            if (!rc && hipMemcpyAsync(x->h_misc, misc, M_COUNT * 4, hipMemcpyDeviceToHost, x->st) != hipSuccess)
                rc = fgss_fail(FGSS_ERR_DEVICE, "copy of the index counters failed");
            if (!rc && hipStreamSynchronize(x->st) != hipSuccess) rc = fgss_fail(FGSS_ERR_DEVICE, "index build failed");
        }
    }
    if (rc) { fgss_server_free(srv); return rc; }
    srv->index_entries = x->h_misc[M_NE];
    srv->mean_list = x->n_terms ? std::max(1.0, (double)x->h_misc[M_NE] / (double)x->n_terms) : 1.0;
    // the self-join workspace of the index is no longer needed: keep postings,
    // list offsets, ids and fences only
    DevBuf *spare[] = {&x->b_ka, &x->b_kb, &x->b_ia, &x->b_ib, &x->b_sid_of, &x->b_pl_of, &x->b_sortscr, &x->b_rng,
                       &x->b_items, &x->b_ovf, &x->b_bounds, &x->b_out};
    for (DevBuf *b : spare)
        if (b->p) { (void)hipFree(b->p); b->p = nullptr; b->cap = 0; }
    srv->slots.resize((size_t)slots);
    for (ServeSlot &sl : srv->slots) {
        sl.d = new fgss_device_sets();
        sl.d->device = x->device;
        bool ok = hipStreamCreateWithFlags(&sl.d->st, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
                  hipHostMalloc((void **)&sl.d->h_ctr, C_COUNT * sizeof(uint64_t), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void **)&sl.d->h_misc, 16 * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
        for (auto &e : sl.d->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
        if (!ok) {
            fgss_server_free(srv);
            return fgss_fail(FGSS_ERR_DEVICE, "server slot set-up failed");
        }
        // workspace for runs of up to max_probes probes of the index's mean length
        // (at most 2^24 tokens) and kReserveJobs requests
        if (!srv->empty) {
            const int64_t mean_len = (x->ntok + x->n - 1) / x->n;
            // work items: as many as the budget's bytes would hold (48 B each), plus the two
            // half-probes of every probe (about 5.7M items, 270 MB, at the default 256 MiB)
            const int64_t items = srv->budget / (int64_t)sizeof(Item) + 2 * srv->max_probes + 4096;
            rc = slot_reserve(sl, srv->max_probes, std::min<int64_t>(srv->max_probes * mean_len, 1ll << 24), kReserveJobs,
                              std::min<int64_t>(items, 1ll << 23));
            if (rc) { fgss_server_free(srv); return rc; }
        }
    }
    *out = srv;
    return FGSS_OK;
}

extern "C" int fgss_server_submit(fgss_server *srv, const fgss_collection *probes, float t, int64_t *request_id)
{
    if (!srv || !probes || !request_id) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (!(t >= srv->t_min) || !(t < 1.0f))
        return fgss_fail(FGSS_ERR_ARG, "threshold %g outside [t_min, 1.0) of the server", (double)t);
    if (!probes->jend.empty()) return fgss_fail(FGSS_ERR_ARG, "probes must be a plain collection");
    if ((int64_t)probes->len.size() > srv->max_probes)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "more than %lld probes in one request", (long long)srv->max_probes);
    for (size_t i = 0; i < probes->len.size(); i++) {
        if (probes->len[i] >= 65536) return fgss_fail(FGSS_ERR_UNSUPPORTED, "probe %zu longer than 65535 tokens", i);
        const int32_t *r = probes->tok.data() + probes->pos[i];
        for (int32_t k = 0; k < probes->len[i]; k++)
            if (r[k] < 0 || (k > 0 && r[k] <= r[k - 1]))
                return fgss_fail(FGSS_ERR_UNSUPPORTED, "probe %zu: tokens must be non-negative and strictly ascending", i);
    }
    auto *r = new ServeRequest();
    r->id = srv->next_id++;
    r->probes = probes;
    r->t = t;
    r->est = serve_request_bytes(probes, t, srv->mean_list, srv->cfg.target);
    r->t_submit = std::chrono::steady_clock::now();
    srv->reqs[r->id] = r;
    srv->pending.push_back(r);
    *request_id = r->id;
    return server_progress(srv, false);
}

extern "C" int fgss_server_poll(fgss_server *srv, int32_t block, int32_t *pending, int32_t *running)
{
    if (!srv) return fgss_fail(FGSS_ERR_ARG, "null server");
    int rc = server_progress(srv, block != 0);
    if (pending) *pending = (int32_t)srv->pending.size();
    if (running) {
        int32_t k = 0;
        for (ServeSlot &sl : srv->slots) k += sl.busy ? (int32_t)sl.reqs.size() : 0;
        *running = k;
    }
    return rc;
}

static ServeRequest *find_request(fgss_server *srv, int64_t id)
{
    auto it = srv->reqs.find(id);
    return it == srv->reqs.end() ? nullptr : it->second;
}

extern "C" int fgss_server_done(fgss_server *srv, int64_t request_id, int32_t *done)
{
    if (!srv || !done) return fgss_fail(FGSS_ERR_ARG, "null argument");
    ServeRequest *r = find_request(srv, request_id);
    if (!r) return fgss_fail(FGSS_ERR_ARG, "unknown request %lld", (long long)request_id);
    *done = r->done ? 1 : 0;
    return FGSS_OK;
}

extern "C" int fgss_server_wait(fgss_server *srv, int64_t request_id)
{
    if (!srv) return fgss_fail(FGSS_ERR_ARG, "null server");
    ServeRequest *r = find_request(srv, request_id);
    if (!r) return fgss_fail(FGSS_ERR_ARG, "unknown request %lld", (long long)request_id);
    while (!r->done) {
        bool busy = false;
        for (ServeSlot &sl : srv->slots) busy |= sl.busy;
        if (!busy && srv->hold)
            return fgss_fail(FGSS_ERR_ARG, "request %lld is queued while admission is held", (long long)request_id);
        int rc = server_progress(srv, true);
        if (rc && !r->done) return rc;
    }
    return FGSS_OK;
}

extern "C" int fgss_server_hold(fgss_server *srv, int32_t hold)
{
    if (!srv) return fgss_fail(FGSS_ERR_ARG, "null server");
    srv->hold = hold != 0;
    return hold ? FGSS_OK : server_progress(srv, false);
}

extern "C" int fgss_server_result(fgss_server *srv, int64_t request_id, fgss_pairs **out, fgss_request_stats *stats)
{
    if (!srv || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    *out = nullptr;
    ServeRequest *r = find_request(srv, request_id);
    if (!r) return fgss_fail(FGSS_ERR_ARG, "unknown request %lld", (long long)request_id);
    if (!r->done) return fgss_fail(FGSS_ERR_ARG, "request %lld is still running", (long long)request_id);
    srv->reqs.erase(request_id);
    std::unique_ptr<ServeRequest> own(r);
    if (r->status != FGSS_OK) return fgss_fail(r->status, "request %lld failed on the device", (long long)request_id);
    std::vector<uint64_t> order;
    sort_pairs_by_key(r->tri, order);
    const size_t R = order.size(), nb = std::max<size_t>(R, 1);
    auto *p = (fgss_pairs *)calloc(1, sizeof(fgss_pairs));
    if (p) {
        p->q = (int32_t *)malloc(nb * 4); p->s = (int32_t *)malloc(nb * 4); p->rid = (int32_t *)malloc(nb * 4);
        p->sid = (int32_t *)malloc(nb * 4); p->overlap = (int32_t *)malloc(nb * 4);
        p->similarity = (float *)malloc(nb * 4);
    }
    if (!p || !p->q || !p->s || !p->rid || !p->sid || !p->overlap || !p->similarity) {
        fgss_pairs_free(p);
        return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    }
    const fgss_collection *Rc = r->probes, *Sc = srv->coll;
    for (size_t k = 0; k < R; k++) {
        const size_t i = (size_t)(order[k] & 0xFFFFFFFFull);
        const int32_t q = r->tri[3 * i], s = r->tri[3 * i + 1], ov = r->tri[3 * i + 2];
        p->q[k] = q;
        p->s[k] = s;
        p->rid[k] = Rc->id[(size_t)q];
        p->sid[k] = Sc->id[(size_t)s];
        p->overlap[k] = ov;
        p->similarity[k] = fgss_similarity((uint32_t)Rc->len[(size_t)q], (uint32_t)Sc->len[(size_t)s], (uint32_t)ov);
    }
    p->n_pairs = (int64_t)R;
    for (int32_t l : Rc->len)
        p->stats.probe_tokens += (int64_t)std::min<uint32_t>(fgss_max_prefix((uint32_t)l, r->t), (uint32_t)l);
    p->stats.index_entries = (int64_t)srv->index_entries;
    p->stats.postings_scanned = (int64_t)r->cnt[0];
    p->stats.candidates = (int64_t)r->cnt[1];
    p->stats.intersections = (int64_t)r->cnt[2];
    p->stats.results = (int64_t)r->cnt[3];
    p->stats.kernel_launches = r->launches;
    if (stats) {
        stats->ms_queue = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
        stats->ms_latency = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        stats->batch_requests = r->batch_requests;
        stats->charged_bytes = r->est;
        stats->inflight_bytes = r->inflight;
        stats->reserved = 0;
    }
    *out = p;
    return FGSS_OK;
}

extern "C" void fgss_server_free(fgss_server *srv)
{
    if (!srv) return;
    for (ServeSlot &sl : srv->slots) {
        if (!sl.d) continue;
        if (sl.d->st) (void)hipStreamSynchronize(sl.d->st);
        DevBuf *bufs[] = {&sl.b_in, &sl.b_sig, &sl.b_sig64, &sl.b_sigbad, &sl.b_wlo};
        for (DevBuf *b : bufs)
            if (b->p) (void)hipFree(b->p);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        sl.h_in = nullptr;
        // the record arrays belong to the slot buffers freed above
        sl.d->tok = sl.d->len = sl.d->pos = sl.d->sig = sl.d->sig64 = sl.d->job = sl.d->jend = nullptr;
        sl.d->sigbad = nullptr;
        sl.d->tj = nullptr;
        if (sl.done) (void)hipEventDestroy(sl.done);
        fgss_collection holder;
        holder.dev = sl.d;
        fgss_device_release(&holder);
    }
    for (auto &kv : srv->reqs) delete kv.second;
    if (srv->idx) {
        fgss_collection holder;
        holder.dev = srv->idx;
        fgss_device_release(&holder);
    }
    delete srv;
}

// ---------------------------------------------------------------------------
// fgss_selfjoin_noalloc: the self-join of records the caller already holds in
// device memory, with every workspace buffer carved out of one static device
// array.  The call makes no device or host allocation (no hipMalloc, hipFree,
// pinned memory, events or streams): for callers that cannot allocate.  It runs
// the same kernels as fgss_selfjoin; capacities are fixed by the array, so a
// collection or a run that does not fit is rejected instead of growing buffers.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxSets = FGSS_NOALLOC_MAX_SETS;
constexpr int64_t kNoAllocMaxTokens = FGSS_NOALLOC_MAX_TOKENS;
constexpr int64_t kNoAllocMaxTerms = FGSS_NOALLOC_MAX_TERMS;
constexpr uint32_t kNoAllocItems = 1u << 18;    // work items of the main launch
constexpr uint32_t kNoAllocBounds = 1u << 19;   // boundary-table entries
constexpr uint32_t kNoAllocSplit = 1u << 16;    // re-split items per launch; overflow records per launch
constexpr uint32_t kNoAllocPending = 1u << 15;  // overflowed items waiting for the next re-split round
constexpr int kNoAllocMaxCUs = 512;             // wavefront counter slots for up to this many CUs
constexpr size_t kNoAllocBytes = (size_t)FGSS_NOALLOC_SCRATCH_BYTES;

// The workspace of every no-allocation call on this GPU (calls that overlap on one
// GPU race on it).  Host code never uses this name as an address: the device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device (looked up on its first use there).
static int noalloc_scratch(int dev, uint8_t **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fgss_fail(FGSS_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fgss_fail(FGSS_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = (uint8_t *)g_scratch_ptr[dev];
    return FGSS_OK;
}

// Carves the workspace of a no-allocation run out of `base` (null: only sizes it),
// the buffers and sizes of prepare_buffers at their upper bounds.  Returns the bytes.
struct NoAllocBufs {
    uint32_t *err = nullptr;     // validation flags
    Item *split = nullptr;       // children of a re-split launch (kNoAllocSplit)
    Item *pend[2] = {nullptr, nullptr};   // overflowed items between rounds (kNoAllocPending each)
};

static size_t noalloc_carve(fgss_device_sets *d, uint8_t *base, int64_t n, int64_t ntok, uint32_t n_terms, int cus,
                            NoAllocBufs *nb)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        bytes = std::max<size_t>(bytes, 16);
        void *p = base ? (void *)(base + off) : nullptr;
        off = (off + bytes + 255) & ~(size_t)255;
        return p;
    };
    auto buf = [&](DevBuf &b, size_t bytes) {
        b.p = take(bytes);
        b.cap = bytes;
    };
    const int64_t N = std::max<int64_t>(n, 1), E = std::max<int64_t>(ntok, 1);
    d->cap_items = kNoAllocItems;
    d->cap_bounds = kNoAllocBounds;
    d->cap_ovf = kNoAllocSplit;
    buf(d->b_midlen, (size_t)N * 4);
    buf(d->b_midoff, (size_t)(N + 1) * 4);
    buf(d->b_probelen, (size_t)N * 4);
    buf(d->b_poff, (size_t)(N + 1) * 4);
    buf(d->b_scanscr, (size_t)scan_scratch_elems(N + 1) * 8);
    buf(d->b_ka, (size_t)E * 4);
    buf(d->b_kb, (size_t)E * 4);
    buf(d->b_ia, (size_t)E * 4);
    buf(d->b_ib, (size_t)E * 4);
    buf(d->b_sid_of, (size_t)E * 4);
    buf(d->b_pl_of, (size_t)E * 4);
    buf(d->b_post, (size_t)(E + kWave) * 8);
    buf(d->b_pden, (size_t)(E + kWave) * 4);
    buf(d->b_psid, (size_t)(E + 1) * 4);
    buf(d->b_fence, (size_t)(E / kFence + 2) * 4);
    buf(d->b_loff, ((size_t)n_terms + 2) * 4);
    buf(d->b_sortscr, (size_t)radix_scratch_elems(E) * 4);
    buf(d->b_rng, (size_t)E * sizeof(uint2));
    buf(d->b_hi, (size_t)N * 4);
    buf(d->b_eq, (size_t)N * 4);
    buf(d->b_nparts, (size_t)N * 4);
    buf(d->b_partoff, (size_t)(N + 1) * 4);
    buf(d->b_bcnt, (size_t)N * 8);
    buf(d->b_bofs, (size_t)(N + 1) * 8);
    buf(d->b_nitems, (size_t)N * 4);
    buf(d->b_dflag, (size_t)N * 4);
    buf(d->b_doff, (size_t)(N + 1) * 4);
    buf(d->b_bounds, (size_t)d->cap_bounds * 4);
    buf(d->b_items, (size_t)d->cap_items * sizeof(Item));
    buf(d->b_dlist, (size_t)N * 4);
    buf(d->b_grp, 8);
    buf(d->b_ovf, (size_t)d->cap_ovf * 4);
    buf(d->b_ovfp, (size_t)d->cap_ovf * 4);
    buf(d->b_wctr, (size_t)std::max(cus, 1) * 2 * 32 * kWaveSlot * 8);   // the two roles of a join launch
    buf(d->b_ctr, C_COUNT * sizeof(uint64_t));
    buf(d->b_jctr, kJobCtr * sizeof(uint64_t));
    buf(d->b_misc, M_COUNT * 4);
    d->sig = (uint32_t *)take((size_t)N * 4 * kSigWords);
    d->sig64 = (uint32_t *)take((size_t)N * 4 * kSig64Words);
    d->sigbad = (uint8_t *)take((size_t)N);
    d->job = (uint32_t *)take((size_t)N * 4);
    d->jend = (uint32_t *)take(4);
    d->tj = (float *)take(4);
    nb->err = (uint32_t *)take(4);
    nb->split = (Item *)take((size_t)kNoAllocSplit * sizeof(Item));
    nb->pend[0] = (Item *)take((size_t)kNoAllocPending * sizeof(Item));
    nb->pend[1] = (Item *)take((size_t)kNoAllocPending * sizeof(Item));
    return off;
}

// Checks of caller-given records before any kernel indexes with them: lengths below
// 2^16, offsets the exclusive scan of the lengths, tokens strictly ascending and in
// [0, n_terms); flag 2 marks lengths that are not non-increasing (allowed: the job is
// then joined in any record order).  A thread per record; flags or-ed into *err.
__global__ void k_validate_records(const uint32_t *__restrict__ tok, const uint32_t *__restrict__ len,
                                   const uint32_t *__restrict__ pos, int64_t n, int64_t ntok, uint32_t n_terms,
                                   uint32_t *__restrict__ err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = len[i], p = pos[i];
    uint32_t e = 0;
    if (l >= 65536u) e |= 1u;
    if (i > 0 && l > len[i - 1]) e |= 2u;
    if ((i == 0 && p != 0u) || (i > 0 && (uint64_t)p != (uint64_t)pos[i - 1] + len[i - 1])) e |= 4u;
    if ((uint64_t)p + l > (uint64_t)ntok || (i == n - 1 && (uint64_t)p + l != (uint64_t)ntok)) e |= 4u;
    if ((e & 4u) == 0) {   // the record's range lies inside the token array
        uint32_t prev = 0;
        for (uint32_t k = 0; k < l; k++) {
            const uint32_t v = tok[p + k];
            if (v >= n_terms) e |= 8u;   // also a negative int32 id
            if (k > 0 && v <= prev) e |= 16u;
            prev = v;
        }
    }
    if (e) atomicOr(err, e);
}

// overflowed item ovf[i] of `in` (narrowed to its overflowed part for a streamed item)
// as a searched item over its source sub-window
__global__ void k_take_ovf(const Item *__restrict__ in, const uint32_t *__restrict__ ovf,
                           const uint32_t *__restrict__ ovfp, uint32_t novf, Item *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= novf) return;
    Item it = in[ovf[i]];
    if (ovfp[i] != kEmpty) {
        const uint64_t W0 = (uint64_t)it.b - it.a;
        const uint32_t a = part_start(it.a, W0, ovfp[i], it.nparts), b = part_start(it.a, W0, ovfp[i] + 1, it.nparts);
        it.a = a;
        it.b = b;
    }
    it.pad &= kPadMask;
    it.part = 0;
    it.nparts = 0;
    it.bofs = 0;
    out[i] = it;
}

// every item of `in` split into `ways` equal source sub-windows
__global__ void k_split_items(const Item *__restrict__ in, uint32_t n, uint32_t ways, Item *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Item it = in[i];
    const uint64_t W = (uint64_t)it.b - it.a;
    for (uint32_t p = 0; p < ways; p++) {
        Item o = it;
        o.a = (uint32_t)(it.a + W * p / ways);
        o.b = (uint32_t)(it.a + W * (p + 1) / ways);
        out[i * ways + p] = o;
    }
}

extern "C" int fgss_selfjoin_noalloc(const int32_t *tokens, const int32_t *lengths, const int32_t *offsets,
                                     int64_t n_sets, int64_t n_tokens, int32_t n_terms, float threshold,
                                     const fgss_join_opts *opts, int32_t *out, int64_t out_cap, int64_t *n_out,
                                     fgss_join_stats *stats, void *stream)
{
    // argument and size checks first: no device call before them
    if (!n_out) return fgss_fail(FGSS_ERR_ARG, "null n_out");
    *n_out = 0;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n_sets < 0 || n_tokens < 0 || n_terms < 0 || out_cap < 0) return fgss_fail(FGSS_ERR_ARG, "negative size");
    if (n_sets > 0 && (!lengths || !offsets || (n_tokens > 0 && !tokens)))
        return fgss_fail(FGSS_ERR_ARG, "null record array");
    if (out_cap > 0 && !out) return fgss_fail(FGSS_ERR_ARG, "null output buffer");
    if (!(threshold > 0.0f) || !(threshold < 1.0f))
        return fgss_fail(FGSS_ERR_ARG, "Threshold must be a float in the range (0.0..1.0).");
    if (n_sets > kNoAllocMaxSets || n_tokens > kNoAllocMaxTokens || n_terms > kNoAllocMaxTerms)
        return fgss_fail(FGSS_ERR_UNSUPPORTED,
                         "%lld records, %lld tokens, %d terms: the no-allocation scratch holds at most %lld records, "
                         "%lld tokens and %lld terms",
                         (long long)n_sets, (long long)n_tokens, n_terms, (long long)kNoAllocMaxSets,
                         (long long)kNoAllocMaxTokens, (long long)kNoAllocMaxTerms);
    const int ts = opts && opts->table_slots ? opts->table_slots : 512;
    if (ts != 256 && ts != 512 && ts != 1024 && ts != 2048 && ts != 4096)
        return fgss_fail(FGSS_ERR_ARG, "table_slots must be 256, 512, 1024, 2048 or 4096");
    if (n_sets == 0) return FGSS_OK;
    auto t0 = std::chrono::steady_clock::now();
    RunCfg cfg = make_cfg(opts);
    cfg.dense_w = 0;   // no dense parts: their offset table would not fit the scratch
    const hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 0;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus > kNoAllocMaxCUs) return fgss_fail(FGSS_ERR_UNSUPPORTED, "more than %d compute units", kNoAllocMaxCUs);
    uint8_t *base = nullptr;
    int rc = noalloc_scratch(dev, &base);
    if (rc) return rc;
    fgss_device_sets d;   // on the stack: its buffers point into the scratch
    NoAllocBufs nb;
    if (noalloc_carve(&d, nullptr, n_sets, n_tokens, (uint32_t)n_terms, cus, &nb) > kNoAllocBytes)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "workspace exceeds the %zu-byte no-allocation scratch", kNoAllocBytes);
    noalloc_carve(&d, base, n_sets, n_tokens, (uint32_t)n_terms, cus, &nb);
    rc = check_range(opts, n_sets, true);
    if (rc) return rc;
    if (opts) {
        d.q_begin = (int32_t)opts->query_begin;
        d.q_end = (int32_t)opts->query_end;
    }
    d.device = dev;
    d.st = st;
    d.n = n_sets;
    d.ntok = n_tokens;
    d.n_terms = (uint32_t)n_terms;
    d.njobs = 1;
    d.max_len = 65535;
    d.tok = (uint32_t *)(n_tokens > 0 ? tokens : lengths);   // any readable address when there are no tokens
    d.len = (uint32_t *)lengths;
    d.pos = (uint32_t *)offsets;
    d.b_out.p = out;
    d.b_out.cap = (size_t)out_cap * 12;
    d.cap_out = (uint32_t)std::min<int64_t>(out_cap, UINT32_MAX / 16);
    uint64_t h_ctr[C_COUNT] = {0}, h_jctr[kJobCtr] = {0};
    uint32_t h_misc[M_COUNT] = {0}, h_err = 0;
    d.h_ctr = h_ctr;
    d.h_misc = h_misc;
    d.h_jctr = h_jctr;
    d.out_hint = 0;   // no pinned result mirror

    // 1. validate the caller's records
    HIPCK(fill<uint32_t>(nb.err, 1, 0u, st));
    hipLaunchKernelGGL(k_validate_records, dim3(grid_for(n_sets, 256)), dim3(256), 0, st, d.tok, d.len, d.pos, n_sets,
                       n_tokens, (uint32_t)n_terms, nb.err);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(&h_err, nb.err, 4, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (h_err & 1u) return fgss_fail(FGSS_ERR_UNSUPPORTED, "records longer than 65535 tokens");
    if (h_err & 4u) return fgss_fail(FGSS_ERR_ARG, "offsets are not the exclusive scan of the lengths");
    if (h_err & 8u) return fgss_fail(FGSS_ERR_UNSUPPORTED, "token ids must lie in [0, n_terms)");
    if (h_err & 16u) return fgss_fail(FGSS_ERR_UNSUPPORTED, "tokens must be strictly ascending in each record");

    // 2. the record set's job table, threshold and signatures, then the pipeline
    HIPCK(fill<uint32_t>(d.job, n_sets, 0u, st));
    HIPCK(fill<uint32_t>(d.jend, 1, (uint32_t)n_sets | ((h_err & 2u) ? kJobUnsorted : 0u), st));
    HIPCK(fill<uint32_t>(P<uint32_t>(d.b_grp), 2, 0u, st));   // one job group: every record
    HIPCK(fill<float>(d.tj, 1, threshold, st));
    hipLaunchKernelGGL(k_signature, dim3((unsigned)std::min<int64_t>((n_sets + 3) / 4, 65536)), dim3(256), 0, st, d.tok,
                       d.pos, d.len, n_sets, d.sig, d.sig64, d.sigbad);
    HIPCK(hipGetLastError());
    rc = enqueue_pipeline(&d, cfg, false);
    if (rc) return rc;
    HIPCK(hipStreamSynchronize(st));
    if (h_misc[M_NITEMS] > d.cap_items)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "%u work items exceed the no-allocation scratch's %u", h_misc[M_NITEMS],
                         d.cap_items);
    if (h_ctr[C_OVERFLOW] > kNoAllocPending)
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "%llu overflowed work items exceed the no-allocation scratch's %u",
                         (unsigned long long)h_ctr[C_OVERFLOW], kNoAllocPending);
    uint64_t written = h_ctr[C_RESULTS];
    if (written > (uint64_t)out_cap) {
        *n_out = (int64_t)written;
        return fgss_fail(FGSS_ERR_ARG, "the output holds %lld triples, the join found at least %llu", (long long)out_cap,
                         (unsigned long long)written);
    }
    uint64_t acc[C_COUNT] = {0};
    for (int k = C_ESCAN; k <= C_DPOST; k++) acc[k] += h_ctr[k];

    // 3. overflowed work items: split four ways and run again, in launches of at most
    //    kNoAllocSplit items; the items that overflow again wait in the other buffer
    uint32_t npend = (uint32_t)h_ctr[C_OVERFLOW];
    int cur = 0, launches = 1;
    if (npend) {
        hipLaunchKernelGGL(k_take_ovf, dim3(grid_for(npend, 256)), dim3(256), 0, st, P<Item>(d.b_items),
                           P<uint32_t>(d.b_ovf), P<uint32_t>(d.b_ovfp), npend, nb.pend[0]);
        HIPCK(hipGetLastError());
    }
    constexpr uint32_t ways = 4, chunk = kNoAllocSplit / ways;
    for (int round = 0; npend > 0; round++) {
        if (round >= 48) return fgss_fail(FGSS_ERR_DEVICE, "internal: work-item splitting did not converge");
        uint32_t nnext = 0;
        for (uint32_t c0 = 0; c0 < npend; c0 += chunk) {
            const uint32_t nc = std::min(chunk, npend - c0), cnt = nc * ways;
            hipLaunchKernelGGL(k_split_items, dim3(grid_for(nc, 256)), dim3(256), 0, st, nb.pend[cur] + c0, nc, ways,
                               nb.split);
            HIPCK(hipGetLastError());
            uint32_t *misc = P<uint32_t>(d.b_misc);
            HIPCK(fill<unsigned long long>(P<unsigned long long>(d.b_ctr), C_COUNT, 0ull, st));
            HIPCK(fill<uint32_t>(misc + M_NRESPLIT, 1, cnt, st));
            JoinArgs A = join_args(&d);
            A.items = nb.split;
            A.n_items_dev = misc + M_NRESPLIT;
            A.n_items_cap = cnt;
            A.out = out + 3 * written;
            A.out_cap = (uint32_t)std::min<uint64_t>((uint64_t)out_cap - written, UINT32_MAX / 16);
            A.ovf_cap = cnt;
            rc = launch_join_cap(&d, A, cfg.cap, false, false, false);   // re-split items are never dense
            if (rc) return rc;
            launches++;
            HIPCK(hipMemcpyAsync(h_ctr, d.b_ctr.p, C_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
            HIPCK(hipStreamSynchronize(st));
            if (written + h_ctr[C_RESULTS] > (uint64_t)out_cap) {
                *n_out = (int64_t)(written + h_ctr[C_RESULTS]);
                return fgss_fail(FGSS_ERR_ARG, "the output holds %lld triples, the join found at least %lld",
                                 (long long)out_cap, (long long)*n_out);
            }
            written += h_ctr[C_RESULTS];
            for (int k = C_ESCAN; k <= C_DPOST; k++) acc[k] += h_ctr[k];
            const uint32_t o = (uint32_t)h_ctr[C_OVERFLOW];
            if ((uint64_t)nnext + o > kNoAllocPending)
                return fgss_fail(FGSS_ERR_UNSUPPORTED, "overflowed work items exceed the no-allocation scratch's %u",
                                 kNoAllocPending);
            if (o) {
                hipLaunchKernelGGL(k_take_ovf, dim3(grid_for(o, 256)), dim3(256), 0, st, nb.split, P<uint32_t>(d.b_ovf),
                                   P<uint32_t>(d.b_ovfp), o, nb.pend[cur ^ 1] + nnext);
                HIPCK(hipGetLastError());
            }
            nnext += o;
        }
        cur ^= 1;
        npend = nnext;
    }
    HIPCK(hipMemcpyAsync(h_jctr, d.b_jctr.p, kJobCtr * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    *n_out = (int64_t)written;
    if (stats) {   // the counters fgss_selfjoin reports (no timing events here)
        stats->index_entries = h_misc[M_NE];
        stats->probe_tokens = h_misc[M_NP];
        stats->postings_scanned = (int64_t)h_jctr[0];
        stats->candidates = (int64_t)h_jctr[1];
        stats->intersections = (int64_t)h_jctr[2];
        stats->results = (int64_t)h_jctr[3];
        stats->verify_tokens = (int64_t)acc[C_VTOK];
        stats->signature_checks = (int64_t)acc[C_SIGCHK];
        stats->signature_checks_128 = (int64_t)acc[C_SIG2];
        stats->verified = (int64_t)acc[C_VERIFIED];
        for (int k = 0; k < 6; k++) stats->phase_cycles[k] = (int64_t)acc[C_CYC_SETUP + k];
        stats->window_retries = (int64_t)acc[C_RETRY];
        stats->dense_parts = (int64_t)acc[C_DENSE];
        stats->dense_postings = (int64_t)acc[C_DPOST];
        stats->kernel_launches = launches;
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return FGSS_OK;
}

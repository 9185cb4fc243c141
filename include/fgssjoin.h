/*
 * fgssjoin.h -- C ABI of the MI355X-native fgssjoin set-similarity self-join.
 *
 * The reference (RafaelQuirino/new_fgssjoin) has no library: its "operator"
 * boundary is four C++ host functions called by its CLI main
 * (src/fgssjoin_cli.cu:61,86,91,100).  Each entry point below names the
 * reference interface it replaces.  Plain pointers and sizes only; every call
 * returns an fgss_status and never exits the process (the reference calls
 * exit(1) from gpuAssert, src/util.cu:43-49, and from its loader).
 *
 * Data model.  A collection is a list of records; each record is a set of
 * int32 token ids (ascending, unique).  Internally records are kept in the
 * reference's order: the .bin file order reversed (src/fgssjoin/data.cu:99-112),
 * i.e. descending length for files written by the q-gram tokenizer.  A result
 * pair (q, s) always has internal index q < s and is printed as
 * "id[q] id[s] similarity" (src/fgssjoin/block.cu:193).
 *
 * Threading.  Handles are thread-compatible: one thread per handle at a time.
 * All device work runs on the current HIP device, on a stream the handle owns
 * or on the one fgss_join_opts.stream names; no call waits for the whole device
 * (buffers a handle outgrows are freed with the handle, since hipFree would).
 */
#ifndef FGSSJOIN_H
#define FGSSJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FGSS_OK = 0,
    FGSS_ERR_ARG = 1,         /* invalid argument: null, t outside (0,1), q < 1 */
    FGSS_ERR_IO = 2,          /* file missing or unreadable ("File error")      */
    FGSS_ERR_FORMAT = 3,      /* truncated or malformed .bin ("Reading error.") */
    FGSS_ERR_DEVICE = 4,      /* a HIP runtime call failed                      */
    FGSS_ERR_NOMEM = 5,       /* host or device allocation failed               */
    FGSS_ERR_UNSUPPORTED = 6  /* collection breaks a documented limit           */
} fgss_status;

/* Message of the last failed call on this thread ("" if none). */
const char *fgss_last_error(void);
/* Library version string. */
const char *fgss_version(void);

/* ------------------------------------------------------------------ */
/* Collections (replaces sets_t / ppjoin_format / prepare_data)       */
/* ------------------------------------------------------------------ */
typedef struct fgss_collection fgss_collection;

typedef struct {
    int64_t n_sets;        /* "# Records"  (src/fgssjoin_cli.cu:62)  */
    int64_t n_tokens;      /* "# Tokens"   (src/fgssjoin_cli.cu:64)  */
    int32_t n_terms;       /* "# Terms" = max token + 1 (data.cu:77-78) */
    int32_t max_len;       /* longest record                          */
    double average_size;   /* "# Average size" (data.cu:100-104)      */
    int32_t sorted_desc;   /* 1 if lengths are non-increasing in internal order */
    int32_t on_device;     /* 1 once fgss_collection_upload has run   */
} fgss_collection_info;

/* Load a reference ".bin" file: per record [int32 id][int32 len][len x int32].
 * Replaces ppjoin_format (include/fgssjoin/data.cuh:65-66, src/fgssjoin/data.cu:31-126). */
int fgss_collection_from_bin(const char *path, fgss_collection **out);
/* Same, from an in-memory int32 stream of n_ints values. */
int fgss_collection_from_bin_buffer(const int32_t *ints, int64_t n_ints, fgss_collection **out);

/* Tokenize a text file (one record per non-empty line) into q-gram sets with
 * global-frequency token ids, as the README's "-f <file> -q <q>" front end
 * (README.md:12-15) and the preprocess library (src/preprocess/string/qgram.cpp:20-49,
 * src/preprocess_cli.cpp:266-340) describe.  Record ids are 1-based line numbers
 * among non-empty lines; lines shorter than q chars yield no record. */
int fgss_collection_from_text(const char *path, int q, fgss_collection **out);
int fgss_collection_from_text_buffer(const char *text, int64_t nbytes, int q, fgss_collection **out);
/* Tokenize text into the token-id space of dict (a collection built from text):
 * q-grams dict knows take its ids, others get new ids from dict's n_terms on (in
 * order of first occurrence).  Used for probe records of R-S joins. */
int fgss_collection_from_text_dict(const char *text, int64_t nbytes, const fgss_collection *dict,
                                   fgss_collection **out);

/* The same tokenization on the current HIP device (newline scan, q-gram keys,
 * occurrence numbers, sdbm hashes, frequency ranks and record order from stable
 * radix sorts): results identical to fgss_collection_from_text_buffer, for
 * q <= 8 and texts under 4 GiB (FGSS_ERR_UNSUPPORTED otherwise). */
int fgss_collection_from_text_device(const char *text, int64_t nbytes, int q, fgss_collection **out);

/* Write the collection as a reference .bin (records by size ascending, i.e.
 * internal order reversed) and, when loc_path is non-null and the collection
 * came from text, the .loc stream (q-gram start offset per stored token). */
int fgss_collection_write_bin(const fgss_collection *c, const char *path, const char *loc_path);

int fgss_collection_get_info(const fgss_collection *c, fgss_collection_info *info);
/* Host views of the internal-order arrays (valid until the collection is freed). */
int fgss_collection_arrays(const fgss_collection *c, const int32_t **id, const int32_t **len,
                           const int64_t **pos, const int32_t **tokens);

/* Copy the collection to HBM on the current device and keep it resident.
 * Replaces the H2D part of prepare_data (src/fgssjoin/data.cu:133-153). */
int fgss_collection_upload(fgss_collection *c);

void fgss_collection_free(fgss_collection *c);

/* Block until all device work issued by this process has finished (the
 * reference's cudaDeviceSynchronize after each launch, src/fgssjoin/filtering.cu:38). */
int fgss_device_synchronize(void);

/* ------------------------------------------------------------------ */
/* Self-join (replaces inverted_index + process_blocks)               */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t index_entries;     /* "# Inverted Index Entries" (fgssjoin_cli.cu:92)    */
    int64_t probe_tokens;      /* sum of max_prefix(|q|)                              */
    int64_t postings_scanned;  /* postings with s > q inside the size window          */
    int64_t candidates;        /* "# Candidates"    (block.cu:209)                    */
    int64_t intersections;     /* "# Intersections" (block.cu:210)                    */
    int64_t results;           /* "# Results"       (block.cu:211)                    */
    int64_t verify_tokens;     /* token reads of suffix-merge verification            */
    double ms_index;           /* mid-prefix selection + inverted index build         */
    double ms_join;            /* candidate generation + verification + compaction    */
    double ms_copy;            /* device -> host copy of the result pairs             */
    double ms_total;           /* whole call                                          */
    double ms_kernel;          /* device time of the join kernel launches (HIP events) */
    int64_t kernel_launches;   /* join kernel launches (1 unless buffers grew or work items were re-split) */
    int64_t signature_checks;  /* candidates tested against the 64-bucket overlap bound */
    int64_t signature_checks_128; /* of those, tested again with 128 buckets           */
    int64_t verified;          /* candidates whose tokens were merged                  */
    int64_t phase_cycles[6];   /* join kernel shader cycles summed over wavefronts: item setup,
                                  candidate generation, positional filter, signature bound,
                                  token verification, and the part of candidate generation
                                  spent locating posting ranges (zeros unless the library is
                                  built with -DFGSS_PHASE_CLOCKS=1)                    */
    int64_t window_retries;    /* streamed windows of several parts whose table overflowed and that
                                  ran again part by part (streamed items join a growing number of
                                  consecutive parts at a time while the table stays sparse)  */
    int64_t dense_parts;       /* dense parts joined (windows full enough for a table indexed by the
                                  source; see fgss_join_opts.dense_pct)                        */
    int64_t dense_postings;    /* of postings_scanned, those read by dense parts (4 bytes each;
                                  the others are 8-byte postings)                              */
} fgss_join_stats;

typedef struct {
    int32_t load_pct;          /* postings per work item as % of its table slots, 0 = default (65) */
    int32_t table_slots;       /* LDS candidate table slots per work item: 256..4096, 0 = default (512) */
    int32_t stream_parts;      /* a query split into more parts than this is walked part after part
                                  by wavefronts holding one posting cursor per probe token (32 parts
                                  each); 0 = default: only queries whose boundary rows do not fit
                                  the boundary table                                   */
    int32_t bounds_cap;        /* precomputed split boundaries at most, 0 = default (2^28); the
                                  parts of queries beyond it locate their posting ranges by search */
    int32_t dense_pct;         /* queries whose source window holds at least this many postings per
                                  100 sources (over several dense parts of 2 x table_slots sources)
                                  are joined part after part in a table indexed by the source; 0 =
                                  default (25), < 0 = never                                      */
    int64_t query_begin;       /* self-joins: only the queries q in [query_begin, query_end) (internal
                                  indices), each against every s > q; 0, 0 = all (another empty range:
                                  none).  One GPU's share of
                                  a join whose queries are split over several (the reference's query
                                  tiles, src/fgssjoin/block.cu:109, taken as a contiguous range) */
    int64_t query_end;
    void *stream;              /* hipStream_t carrying the call's device work (the reference's
                                  per-thread default stream, include/util.cuh:4); null = the handle's
                                  own stream.  The call still returns after that stream has drained. */
    int32_t index_groups;      /* batches: the jobs split into this many groups of consecutive jobs
                                  (balanced by records), each indexed and joined by a pipeline run of
                                  its own -- a smaller index, item list and working set per run (the
                                  reference partitions its index by record block instead,
                                  src/fgssjoin/block.cu:273-434, index.cu:114-162); 0 or 1 = one */
    int32_t block_records;     /* a work item spans at most this many consecutive source records
                                  (the reference's record-block size, src/fgssjoin/block.cu:50-81,
                                  as an upper bound on a part's window); 0 = no limit */
} fgss_join_opts;

typedef struct {
    int64_t n_pairs;
    int32_t *q;             /* internal index of the query record (q < s)      */
    int32_t *s;             /* internal index of the source record             */
    int32_t *rid;           /* id[q] as printed                                */
    int32_t *sid;           /* id[s] as printed                                */
    int32_t *overlap;       /* |Q intersect S|                                 */
    float *similarity;      /* (float)ov / (float)(|Q|+|S|-ov)                  */
    fgss_join_stats stats;
} fgss_pairs;

/* Self-join at Jaccard threshold t in (0,1).  The collection is uploaded on
 * first use.  Pairs are returned sorted by (q, s).  Replaces inverted_index
 * (include/fgssjoin/index.cuh:99-100) + process_blocks
 * (include/fgssjoin/block.cuh:13-14), whose results were printed instead of
 * returned.  opts may be null.  Records may come in any order (as ppjoin_format
 * reads them): the earlier internal index of a pair is its query, and a collection
 * whose lengths are not non-increasing joins every query against every later record
 * through the 4-way size filter (slower than a size-ascending .bin, same rules). */
int fgss_selfjoin(fgss_collection *c, float threshold, const fgss_join_opts *opts, fgss_pairs **out);
void fgss_pairs_free(fgss_pairs *p);

/* Self-join without allocating.  No counterpart in the reference: for callers that
 * cannot allocate device or host memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The records are already in
 * device memory in the layout fgss_collection_arrays describes (internal order,
 * any length order, tokens strictly ascending in [0, n_terms), offsets the
 * exclusive scan of the lengths; int32 each); every workspace buffer lies in one
 * static device array of FGSS_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, so collections above the limits below fail with FGSS_ERR_UNSUPPORTED.
 * Result triples (q, s, overlap), the same set fgss_selfjoin returns, in no fixed
 * order, go to the device buffer out (out_cap triples); *n_out receives their
 * number (with FGSS_ERR_ARG and the count found so far when out_cap is too small).
 * stats (may be null) receives the counters, without timing events.  All work runs
 * on `stream` (a hipStream_t; null is the null stream) and the call returns after
 * it has drained.  Calls that overlap on one GPU share its array and race. */
#define FGSS_NOALLOC_MAX_SETS 65536
#define FGSS_NOALLOC_MAX_TOKENS 524288
#define FGSS_NOALLOC_MAX_TERMS 524288
#define FGSS_NOALLOC_SCRATCH_BYTES (72u << 20)
int fgss_selfjoin_noalloc(const int32_t *tokens, const int32_t *lengths, const int32_t *offsets, int64_t n_sets,
                          int64_t n_tokens, int32_t n_terms, float threshold, const fgss_join_opts *opts,
                          int32_t *out, int64_t out_cap, int64_t *n_out, fgss_join_stats *stats, void *stream);

/* Print the self-join's inverted index at threshold t as the reference's debug
 * mode does (print_inverted_index, src/fgssjoin/index.cu:218-243): per term,
 * "term: [t], size: n" and the ids of the records whose mid prefix holds it.  Lists
 * are in internal order (the reference's are in atomic, arbitrary order) and a term
 * with no entries prints its own id (the reference prints a neighbour's). */
int fgss_collection_print_index(fgss_collection *c, float threshold, void *file /* FILE* */);

/* Print pairs as the reference does: "%d %d %f\n" = id[q] id[s] similarity
 * (src/fgssjoin/block.cu:193). */
int fgss_pairs_print(const fgss_pairs *p, void *file /* FILE* */);

/* ------------------------------------------------------------------ */
/* Batches: many independent self-joins in one pipeline run            */
/* (no reference counterpart; the reference runs one join per process) */
/* ------------------------------------------------------------------ */
typedef struct fgss_batch fgss_batch;

typedef struct {
    double ms_total;          /* whole fgss_batch_run call                          */
    double ms_pipeline;       /* device pipeline including its final synchronization */
    int64_t n_pairs;          /* result pairs over all jobs                          */
    int64_t candidates;       /* "# Candidates" summed over all jobs                 */
    int32_t graph;            /* 1 when the pipeline ran as a replayed HIP graph     */
    int32_t kernel_launches;  /* join kernel launches (1 unless buffers grew or items were re-split) */
} fgss_batch_stats;

/* Create a batch of njobs self-joins: job j joins colls[j] at thresholds[j].
 * The collections are copied (concatenated, each keeping a private token range)
 * and uploaded once; the caller keeps ownership of colls and must keep them alive
 * while it reads results (record ids come from them). */
int fgss_batch_create(fgss_collection *const *colls, const float *thresholds, int32_t njobs, fgss_batch **out);
/* Change the jobs' thresholds for the next run (a captured graph stays valid). */
int fgss_batch_set_thresholds(fgss_batch *b, const float *thresholds);
/* Run every job: one launch of each pipeline stage for the whole batch.  With
 * use_graph the stages are captured once into a HIP graph and replayed. */
int fgss_batch_run(fgss_batch *b, const fgss_join_opts *opts, int32_t use_graph, fgss_batch_stats *stats);
/* Pairs of one job from the last run, exactly as fgss_selfjoin would return them
 * (sorted, indices and ids of that job's collection). */
int fgss_batch_pairs(fgss_batch *b, int32_t job, fgss_pairs **out);
/* Pairs of every job from the last run in one result: job j's pairs, exactly as
 * fgss_batch_pairs(b, j) returns them, are entries [job_first[j], job_first[j+1]);
 * job_first holds njobs + 1 entries.  job_stats (njobs entries, or null) receives
 * each job's counters; out->stats is not filled.  No reference counterpart (the
 * reference joins one collection per run). */
int fgss_batch_pairs_all(fgss_batch *b, fgss_pairs **out, int64_t *job_first, fgss_join_stats *job_stats);
/* Autotuner: graph replays of the batch, median of reps each, over LDS table sizes
 * {512, 1024} x work-item loads {50, 65, 80}%, then, from the fastest, index groups
 * {2, 4, 8} and record blocks {4096, 32768}; returns the fastest options. */
int fgss_batch_autotune(fgss_batch *b, int32_t reps, fgss_join_opts *best, double *best_ms);
/* The autotuner with a tuning cache: cache_path, a text file (created when missing),
 * holds one line per tuned workload, keyed by its jobs, records, tokens, thresholds,
 * device and library version.  A workload found there gets its stored choice without
 * timing (*from_cache = 1; *best_ms the stored time); any other is tuned and its line
 * appended.  No reference counterpart. */
int fgss_batch_autotune_cached(fgss_batch *b, int32_t reps, const char *cache_path, fgss_join_opts *best,
                               double *best_ms, int32_t *from_cache);
void fgss_batch_free(fgss_batch *b);

/* ------------------------------------------------------------------ */
/* Serving: R-S joins against a resident index, continuously batched   */
/* (no reference counterpart; its "dualjoin" mode is only a usage      */
/* string, src/preprocess_cli.cpp:26)                                  */
/* ------------------------------------------------------------------ */
typedef struct fgss_server fgss_server;

typedef struct {
    double ms_queue;          /* submit -> launch of the run that carried it     */
    double ms_latency;        /* submit -> its results were on the host          */
    int32_t batch_requests;   /* requests that shared that run (itself included) */
    int32_t reserved;
    int64_t charged_bytes;    /* its charge against the server's budget          */
    int64_t inflight_bytes;   /* charges in flight, over all slots, once its run
                                 was admitted (its own included)                  */
} fgss_request_stats;

/* Index `indexed` (a plain collection in internal order, lengths non-increasing)
 * once for every threshold >= t_min: each record's first max_prefix(|s|, t_min)
 * tokens.  The server keeps a private device copy; the caller keeps `indexed`
 * alive while results are read (record ids come from it).  `slots` runs may be in
 * flight at once, each on its own HIP stream; `budget_bytes` (<= 0: 256 MiB) caps
 * the device memory of the runs' work buffers in flight: each request is charged
 * the bytes it occupies in a run -- its staged records, the plan arrays and posting
 * ranges of its probes, and its work items with their boundary rows, estimated from
 * the postings its probe prefixes may reach (prefix tokens x the index's mean list
 * length); the candidates themselves are counted in the join kernel's LDS.  Requests
 * are admitted while their charges fit.  Larger budgets batch more requests per run:
 * more throughput, longer queueing.  Each slot's workspace is allocated here, for runs
 * of up to 65,536 probes of the index's mean length and the work items of the whole
 * budget, so that serving does not allocate; larger runs grow it. */
int fgss_server_create(const fgss_collection *indexed, float t_min, int32_t slots, int64_t budget_bytes,
                       const fgss_join_opts *opts, fgss_server **out);
/* Queue an R-S join of `probes` (a plain collection in the index's token space,
 * e.g. from fgss_collection_from_text_dict; any length order) at threshold
 * t in [t_min, 1).  Pending requests are admitted, first come first served, into
 * idle slots, several per run, while their charges fit the budget.  The caller
 * keeps `probes` alive until fgss_server_result. */
int fgss_server_submit(fgss_server *srv, const fgss_collection *probes, float t, int64_t *request_id);
/* Retire completed runs and admit pending requests; with block set, first wait for
 * the oldest run in flight.  Reports the requests still queued and running. */
int fgss_server_poll(fgss_server *srv, int32_t block, int32_t *pending, int32_t *running);
int fgss_server_done(fgss_server *srv, int64_t request_id, int32_t *done);
/* Drive the server until the request is done (an error if it is queued while
 * admission is held and nothing runs). */
int fgss_server_wait(fgss_server *srv, int64_t request_id);
/* Hold (1) or resume (0) admission: held, submitted requests only queue, so a
 * caller can hand over a burst that is then batched as one. */
int fgss_server_hold(fgss_server *srv, int32_t hold);
/* Take a done request's pairs: q indexes `probes`, s the indexed collection
 * (both internal order), sorted by (q, s).  A pair (r, s) is a result iff it is one
 * under the self-join's rules with the longer record as the query (the probe on
 * ties): size filter, a common prefix token (max prefix of the longer record, mid
 * prefix of the shorter), positional filter, |R n S| >= floor(min_overlap) and
 * float32 Jaccard >= t -- the same whatever else ran in the same batch.  Serving a
 * collection against itself thus returns the self-join's pairs for q < s.  The
 * request is then forgotten. */
int fgss_server_result(fgss_server *srv, int64_t request_id, fgss_pairs **out, fgss_request_stats *stats);
void fgss_server_free(fgss_server *srv);

/* ------------------------------------------------------------------ */
/* Synthetic workloads (no reference counterpart; used by the bench)  */
/* ------------------------------------------------------------------ */
typedef enum {
    FGSS_GEN_TITLES = 0,   /* 4-15 Zipf words, 20-120 chars, 10% copies with 1-3 edits  */
    FGSS_GEN_LONG = 1      /* 200-1200 chars of Zipf words, 5% copies with 1-10 edits  */
} fgss_gen_kind;

/* Generate n_records lines of seeded synthetic text ('\n'-terminated) into a
 * malloc'd buffer; release with fgss_free. */
int fgss_generate_text(int kind, int64_t n_records, uint64_t seed, char **text, int64_t *nbytes);
void fgss_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* FGSSJOIN_H */

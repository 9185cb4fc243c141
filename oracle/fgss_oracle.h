/*
 * fgss_oracle.h -- CPU oracle for the fgssjoin set-similarity self-join.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU product is
 * compared against, and the CPU baseline that bench.py times.  Nothing in the
 * product path (new_fgssjoin_amd/, include/) may link, load or call it.
 *
 * It restates, in plain C11, the reference's algorithm:
 *   - bound expressions        include/fgssjoin/similarity.cuh:9-82
 *   - .bin loader (reversal)   src/fgssjoin/data.cu:31-126
 *   - mid-prefix selection     src/fgssjoin/data.cu:241-257
 *   - inverted index           src/fgssjoin/index.cu:169-211, :276-342
 *   - candidate filtering      src/fgssjoin/filtering.cu:76-141
 *   - compaction               src/fgssjoin/compaction.cu:8-17
 *   - verification             src/fgssjoin/checking.cu:93-175
 *   - tile loop + final test   src/fgssjoin/block.cu:61-213
 *   - threshold parsing        src/fgssjoin_cli.cu:31-35
 *   - q-gram tokenizer         src/io.cpp:8-26, src/preprocess/string/data.cpp:36-40,
 *                              src/preprocess/string/qgram.cpp:9-49, src/util.cpp:146-157
 *   - frequency reordering     the rule that reproduces the reference fixture
 *                              test/data/string/small.txt.3gram.bin{,.loc} byte for byte
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_tokenize           fixture small.txt -> .bin/.loc bytes (golden vectors)
 *   orc_bin_load           fixture decode (ids, lengths, reversal) + round trip
 *   bounds (orc_*size ...) hand-derived exact values + an independent float64
 *                          restatement of the correctly rounded float32 formulas
 *   orc_selfjoin_queries   == orc_selfjoin restricted to the query subset (pairs and
 *                          counters summed over the subset)
 *   orc_selfjoin_ex        checks = 1: orc_selfjoin / _queries; checks = 0: the same
 *                          triples and counters (v_reads aside), its verification
 *                          being the one the checks=1 run compares with the closed form
 *   orc_selfjoin           == orc_selfjoin_literal for block sizes {1, 7, N}
 *                          (two independent restatements: closed form vs literal
 *                          tile loop), subset of orc_selfjoin_brute with the
 *                          difference confined to float-boundary pairs, and the
 *                          known answers on the fixture collection at 8 thresholds
 *   orc_rsjoin             subset of orc_rsjoin_brute with the difference confined
 *                          to float-boundary pairs; R-S of a collection with itself
 *                          has a q<s half equal to orc_selfjoin (and so reproduces
 *                          the fixture's known answers), holds every non-empty self
 *                          pair and is symmetric away from float-boundary pairs;
 *                          t_index <= t changes no result
 * No function is left "parity unpinned".
 */
#ifndef FGSS_ORACLE_H
#define FGSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* float32 bound expressions, include/fgssjoin/similarity.cuh:9-82 */
float    orc_min_size(uint32_t len, float t);
float    orc_max_size(uint32_t len, float t);
float    orc_min_overlap(uint32_t a, uint32_t b, float t);
uint32_t orc_max_prefix(uint32_t len, float t);
uint32_t orc_mid_prefix(uint32_t len, float t);
float    orc_similarity(uint32_t a, uint32_t b, uint32_t inter);
/* atof -> float, src/fgssjoin_cli.cu:31 */
float    orc_parse_threshold(const char *s);

/* Tokenize a text buffer into the reference ".bin" int32 stream
 * ([id][len][len x token] per record, records by size ascending) and the
 * ".loc" stream (q-gram start offset per stored token).  Returns 0 on success.
 * Buffers are malloc'd; release with orc_free. */
int  orc_tokenize(const char *text, int64_t nbytes, int q,
                  int32_t **bin_out, int64_t *bin_n,
                  int32_t **loc_out, int64_t *loc_n);
void orc_free(void *p);

/* Count and load a .bin int32 stream into the reference's internal order
 * (file order reversed, src/fgssjoin/data.cu:99-112).  pos[] is int64. */
int orc_bin_count(const int32_t *bin, int64_t nints, int64_t *n_sets,
                  int64_t *n_tokens, int32_t *n_terms);
int orc_bin_load(const int32_t *bin, int64_t nints, int32_t *id, int32_t *len,
                 int64_t *pos, int32_t *tok);

typedef struct {
    int64_t e_idx;          /* index entries: sum mid_prefix(|s|)                 */
    int64_t p_q;            /* probe tokens: sum max_prefix(|q|)                  */
    int64_t e_scan;         /* postings with s>q passing the size filter          */
    int64_t m_reads;        /* token reads of the reference merge loop            */
    int64_t v_reads;        /* token reads of a merge over the suffixes past tau  */
    int64_t candidates;     /* pairs passing rules 1-4 (reference "# Candidates") */
    int64_t intersections;  /* candidates with overlap >= floor(bound)           */
    int64_t results;        /* printed pairs                                     */
} orc_counters;

/* Self-join over a collection in internal order.  Result triples
 * (q, s, overlap) in internal indices, sorted by (q, s), malloc'd into
 * *out (3 * return value int32s).  Returns the number of pairs, or -1.
 *   orc_selfjoin          closed form, nthreads worker threads
 *   orc_selfjoin_literal  the reference's tile loop with a dense int16 score
 *                         matrix of block_size^2 buckets, single thread
 *   orc_selfjoin_brute    every pair q<s with float32 Jaccard >= t (superset) */
int64_t orc_selfjoin(int64_t n_sets, const int32_t *len, const int64_t *pos,
                     const int32_t *tok, int32_t n_terms, float t, int nthreads,
                     int32_t **out, orc_counters *ctr);
/* orc_selfjoin restricted to the given queries (internal indices, each against
 * every s > q): their triples and counters (e_idx is the whole index's). */
int64_t orc_selfjoin_queries(int64_t n_sets, const int32_t *len, const int64_t *pos,
                             const int32_t *tok, int32_t n_terms, float t, const int64_t *queries,
                             int64_t n_queries, int nthreads, int32_t **out, orc_counters *ctr);
/* orc_selfjoin (queries null) or orc_selfjoin_queries, with checks = 0: each
 * candidate is verified by the reference's merge alone (checking.cu:93-175, the
 * score it stores), without the closed-form overlap it is checked against and the
 * suffix read count (v_reads stays 0).  The CPU baseline's timed pass.  secs (may
 * be null) receives the wall seconds of the index build and of the query phase. */
int64_t orc_selfjoin_ex(int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                        int32_t n_terms, float t, int nthreads, const int64_t *queries, int64_t n_queries,
                        int checks, int32_t **out, orc_counters *ctr, double *secs);
int64_t orc_selfjoin_literal(int64_t n_sets, const int32_t *len, const int64_t *pos,
                             const int32_t *tok, int32_t n_terms, float t,
                             int64_t block_size, int32_t **out, orc_counters *ctr);
int64_t orc_selfjoin_brute(int64_t n_sets, const int32_t *len, const int64_t *pos,
                           const int32_t *tok, float t, int nthreads, int32_t **out);

/* R-S join of probe records (any order) against an indexed collection (internal
 * order), both in one token space; tokens >= n_terms occur only in probes.
 * Triples (r, s, overlap) sorted by (r, s).  t_index <= t is the threshold the
 * index prefixes were cut at (a server's t_min); it changes only e_scan.
 *   orc_rsjoin        closed form: the self-join's rules with the longer record of
 *                     each pair as the query (the probe on ties), nthreads workers
 *   orc_rsjoin_brute  every pair with float32 Jaccard >= t (superset), one thread */
int64_t orc_rsjoin(int64_t n_probes, const int32_t *plen, const int64_t *ppos, const int32_t *ptok,
                   int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                   int32_t n_terms, float t, float t_index, int nthreads, int32_t **out,
                   orc_counters *ctr);
int64_t orc_rsjoin_brute(int64_t n_probes, const int32_t *plen, const int64_t *ppos, const int32_t *ptok,
                         int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                         float t, int32_t **out);

#ifdef __cplusplus
}
#endif
#endif

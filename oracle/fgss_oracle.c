/*
 * fgss_oracle.c -- CPU oracle for the fgssjoin set-similarity self-join.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.  The
 * product library never links or calls this file.  Every function cites the
 * reference file:line it restates; the header lists what pins each one.
 *
 * Build: see oracle/Makefile (gcc -O3 -march=native -ffp-contract=off).
 * -ffp-contract=off matters: the bound expressions must round exactly like
 * the reference's separate float32 multiply / divide / add.
 */
#define _GNU_SOURCE
#include "fgss_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Bound expressions: include/fgssjoin/similarity.cuh:9-82                   */
/* ------------------------------------------------------------------------ */

/* similarity.cuh:9-13 */
float orc_min_size(uint32_t len, float t) { return (float)len * t; }
/* similarity.cuh:20-24 */
float orc_max_size(uint32_t len, float t) { return (float)len / t; }
/* similarity.cuh:31-36: t/(1+t) first, then times (float)(a+b) */
float orc_min_overlap(uint32_t a, uint32_t b, float t)
{
    float r = t / (1.0f + t);
    return r * (float)(a + b);
}
/* similarity.cuh:43-47: unsigned - ceil(float) + 1, evaluated in float */
uint32_t orc_max_prefix(uint32_t len, float t)
{
    float v = (float)len - ceilf(orc_min_size(len, t)) + 1.0f;
    return (uint32_t)v;
}
/* similarity.cuh:54-58 */
uint32_t orc_mid_prefix(uint32_t len, float t)
{
    float v = (float)len - ceilf(orc_min_overlap(len, len, t)) + 1.0f;
    return (uint32_t)v;
}
/* similarity.cuh:65-82 */
float orc_similarity(uint32_t a, uint32_t b, uint32_t inter)
{
    uint32_t uni = a + b - inter;
    return (float)inter / (float)uni;
}
/* src/fgssjoin_cli.cu:31 -- threshold = atof(argv[2]) assigned to float */
float orc_parse_threshold(const char *s) { return (float)atof(s); }

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Radix sort of (u64 key, u32 payload), stable                              */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t key; uint32_t val; uint32_t pad; } kv64;

static void radix_sort_kv64(kv64 *a, int64_t n, int key_bytes)
{
    kv64 *tmp = (kv64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(kv64));
    kv64 *src = a, *dst = tmp;
    for (int pass = 0; pass < key_bytes; pass++) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int sh = pass * 8;
        for (int64_t i = 0; i < n; i++) cnt[((src[i].key >> sh) & 255) + 1]++;
        for (int k = 0; k < 256; k++) cnt[k + 1] += cnt[k];
        for (int64_t i = 0; i < n; i++) dst[cnt[(src[i].key >> sh) & 255]++] = src[i];
        kv64 *x = src; src = dst; dst = x;
    }
    if (src != a) memcpy(a, src, (size_t)n * sizeof(kv64));
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Tokenizer: src/io.cpp:8-26 (non-empty lines), src/preprocess/string/      */
/* data.cpp:36-40 (A-Z -> a-z), qgram.cpp:20-49 (q-grams with occurrence     */
/* suffix), qgram.cpp:9-13 + src/util.cpp:146-157 (64-bit sdbm hash).        */
/* Reordering (token id = rank by (document frequency, first occurrence),    */
/* tokens ascending, records by size ascending, id = 1-based line number) is */
/* the rule the fixture small.txt.3gram.bin{,.loc} pins.                     */
/* ------------------------------------------------------------------------ */

/* sdbm over the bytes of gram followed by the decimal digits of occ, stopping
 * at the first NUL byte exactly like a C-string walk (util.cpp:146-157). */
static uint64_t sdbm_gram(const unsigned char *g, int q, int occ)
{
    uint64_t h = 0;
    for (int i = 0; i < q; i++) {
        unsigned c = g[i];
        if (c == 0) return h;
        h = (uint64_t)c + (h << 6) + (h << 16) - h;
    }
    char num[24];
    int nd = snprintf(num, sizeof(num), "%d", occ);
    for (int i = 0; i < nd; i++) {
        unsigned c = (unsigned char)num[i];
        h = (uint64_t)c + (h << 6) + (h << 16) - h;
    }
    return h;
}

static uint64_t fnv_bytes(const unsigned char *p, int n)
{
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < n; i++) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

typedef struct { int64_t start; int64_t len; } line_t;

static int cmp_i64pair(const void *a, const void *b)
{
    const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

int orc_tokenize(const char *text, int64_t nbytes, int q, int32_t **bin_out,
                 int64_t *bin_n, int32_t **loc_out, int64_t *loc_n)
{
    if (q < 1 || !text || nbytes < 0) return -1;
    /* lines: getline splits at '\n'; empty lines are skipped (io.cpp:16-19) */
    int64_t cap = 1024, nl = 0;
    line_t *lines = (line_t *)malloc((size_t)cap * sizeof(line_t));
    int64_t s0 = 0;
    for (int64_t i = 0; i <= nbytes; i++) {
        if (i == nbytes || text[i] == '\n') {
            if (i > s0) {
                if (nl == cap) { cap *= 2; lines = (line_t *)realloc(lines, (size_t)cap * sizeof(line_t)); }
                lines[nl].start = s0; lines[nl].len = i - s0; nl++;
            }
            s0 = i + 1;
        }
    }
    /* gram counts */
    int64_t total = 0;
    int64_t *rec_off = (int64_t *)malloc((size_t)(nl + 1) * sizeof(int64_t));
    for (int64_t r = 0; r < nl; r++) {
        rec_off[r] = total;
        int64_t g = lines[r].len - q + 1;
        total += g > 0 ? g : 0;
    }
    rec_off[nl] = total;
    kv64 *kv = (kv64 *)malloc((size_t)(total > 0 ? total : 1) * sizeof(kv64));
    int32_t *gpos = (int32_t *)malloc((size_t)(total > 0 ? total : 1) * sizeof(int32_t));
    unsigned char *buf = NULL; int64_t bufcap = 0;
    int64_t *tab = NULL; int64_t tabcap = 0;    /* open addressing: gram start, -1 empty */
    int32_t *tabocc = NULL;
    for (int64_t r = 0; r < nl; r++) {
        int64_t L = lines[r].len, ng = rec_off[r + 1] - rec_off[r];
        if (ng <= 0) continue;
        if (L > bufcap) { bufcap = L; buf = (unsigned char *)realloc(buf, (size_t)bufcap); }
        for (int64_t i = 0; i < L; i++) {
            unsigned char c = (unsigned char)text[lines[r].start + i];
            buf[i] = (c >= 'A' && c <= 'Z') ? (unsigned char)(c - ('Z' - 'z')) : c;
        }
        int64_t tc = 16;
        while (tc < 2 * ng) tc *= 2;
        if (tc > tabcap) {
            tabcap = tc;
            tab = (int64_t *)realloc(tab, (size_t)tabcap * sizeof(int64_t));
            tabocc = (int32_t *)realloc(tabocc, (size_t)tabcap * sizeof(int32_t));
        }
        for (int64_t k = 0; k < tc; k++) tab[k] = -1;
        for (int64_t i = 0; i < ng; i++) {
            uint64_t h = fnv_bytes(buf + i, q);
            int64_t slot = (int64_t)(h & (uint64_t)(tc - 1));
            int occ;
            for (;;) {
                if (tab[slot] < 0) { tab[slot] = i; tabocc[slot] = 0; occ = 0; break; }
                if (memcmp(buf + tab[slot], buf + i, (size_t)q) == 0) { occ = ++tabocc[slot]; break; }
                slot = (slot + 1) & (tc - 1);
            }
            int64_t k = rec_off[r] + i;
            kv[k].key = sdbm_gram(buf + i, q, occ);
            kv[k].val = (uint32_t)k;
            gpos[k] = (int32_t)i;
        }
    }
    free(buf); free(tab); free(tabocc);

    /* distinct hashes with document frequency and first occurrence */
    radix_sort_kv64(kv, total, 8);
    int64_t ndist = 0;
    for (int64_t i = 0; i < total; i++)
        if (i == 0 || kv[i].key != kv[i - 1].key) ndist++;
    int64_t *grp = (int64_t *)malloc((size_t)(ndist > 0 ? ndist : 1) * 2 * sizeof(int64_t)); /* (df, first) */
    int64_t *grp_of = (int64_t *)malloc((size_t)(total > 0 ? total : 1) * sizeof(int64_t));
    int64_t gi = -1;
    for (int64_t i = 0; i < total; i++) {
        if (i == 0 || kv[i].key != kv[i - 1].key) {
            gi++;
            grp[2 * gi] = 0;
            grp[2 * gi + 1] = kv[i].val;   /* stable sort: first element has min index */
        }
        grp[2 * gi]++;
        grp_of[kv[i].val] = gi;
    }
    /* rank groups by (df asc, first asc) */
    int64_t *order = (int64_t *)malloc((size_t)(ndist > 0 ? ndist : 1) * 3 * sizeof(int64_t));
    for (int64_t g = 0; g < ndist; g++) {
        order[3 * g] = grp[2 * g]; order[3 * g + 1] = grp[2 * g + 1]; order[3 * g + 2] = g;
    }
    qsort(order, (size_t)ndist, 3 * sizeof(int64_t), cmp_i64pair);
    int32_t *rank = (int32_t *)malloc((size_t)(ndist > 0 ? ndist : 1) * sizeof(int32_t));
    for (int64_t k = 0; k < ndist; k++) rank[order[3 * k + 2]] = (int32_t)k;
    free(order); free(grp); free(kv);

    /* records: tokens ascending (with their offsets); records by size ascending, stable */
    int64_t nrec = 0;
    for (int64_t r = 0; r < nl; r++) if (rec_off[r + 1] > rec_off[r]) nrec++;
    int64_t *rorder = (int64_t *)malloc((size_t)(nrec > 0 ? nrec : 1) * 2 * sizeof(int64_t));
    int64_t k2 = 0;
    for (int64_t r = 0; r < nl; r++)
        if (rec_off[r + 1] > rec_off[r]) { rorder[2 * k2] = rec_off[r + 1] - rec_off[r]; rorder[2 * k2 + 1] = r; k2++; }
    qsort(rorder, (size_t)nrec, 2 * sizeof(int64_t), cmp_i64pair);

    int32_t *bin = (int32_t *)malloc((size_t)(2 * nrec + total + 1) * sizeof(int32_t));
    int32_t *loc = (int32_t *)malloc((size_t)(total + 1) * sizeof(int32_t));
    int64_t *pairs = NULL; int64_t pcap = 0;
    int64_t bo = 0, lo = 0;
    for (int64_t k = 0; k < nrec; k++) {
        int64_t r = rorder[2 * k + 1], n = rec_off[r + 1] - rec_off[r];
        if (n > pcap) { pcap = n; pairs = (int64_t *)realloc(pairs, (size_t)pcap * 2 * sizeof(int64_t)); }
        for (int64_t i = 0; i < n; i++) {
            int64_t f = rec_off[r] + i;
            pairs[2 * i] = rank[grp_of[f]];
            pairs[2 * i + 1] = gpos[f];
        }
        qsort(pairs, (size_t)n, 2 * sizeof(int64_t), cmp_i64pair);
        bin[bo++] = (int32_t)(r + 1);
        bin[bo++] = (int32_t)n;
        for (int64_t i = 0; i < n; i++) { bin[bo++] = (int32_t)pairs[2 * i]; loc[lo++] = (int32_t)pairs[2 * i + 1]; }
    }
    free(pairs); free(rorder); free(rank); free(grp_of); free(gpos); free(rec_off); free(lines);
    *bin_out = bin; *bin_n = bo; *loc_out = loc; *loc_n = lo;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* .bin loader: src/fgssjoin/data.cu:31-126                                  */
/* ------------------------------------------------------------------------ */
int orc_bin_count(const int32_t *bin, int64_t nints, int64_t *n_sets,
                  int64_t *n_tokens, int32_t *n_terms)
{
    int64_t i = 0, ns = 0, nt = 0;
    int32_t terms = 0;
    while (i < nints) {
        if (i + 2 > nints) return -1;                 /* data.cu:63-69 "Reading error." */
        int32_t sz = bin[i + 1];
        if (sz < 0 || i + 2 + sz > nints) return -1;  /* data.cu:72-85 */
        for (int32_t k = 0; k < sz; k++)
            if (bin[i + 2 + k] >= terms) terms = bin[i + 2 + k] + 1;   /* data.cu:77-78 */
        nt += sz; ns++; i += 2 + sz;
    }
    *n_sets = ns; *n_tokens = nt; *n_terms = terms;
    return 0;
}

/* Records are stored in reverse file order (data.cu:99-112), so a size-sorted
 * file yields descending lengths; pos is the exclusive scan of len (:115-118). */
int orc_bin_load(const int32_t *bin, int64_t nints, int32_t *id, int32_t *len,
                 int64_t *pos, int32_t *tok)
{
    int64_t ns, nt; int32_t terms;
    if (orc_bin_count(bin, nints, &ns, &nt, &terms) != 0) return -1;
    int64_t *start = (int64_t *)malloc((size_t)(ns > 0 ? ns : 1) * sizeof(int64_t));
    int64_t i = 0, r = 0;
    while (i < nints) { start[r++] = i; i += 2 + bin[i + 1]; }
    int64_t p = 0;
    for (int64_t k = 0; k < ns; k++) {
        int64_t s = start[ns - 1 - k];
        id[k] = bin[s]; len[k] = bin[s + 1]; pos[k] = p;
        memcpy(tok + p, bin + s + 2, (size_t)bin[s + 1] * sizeof(int32_t));
        p += bin[s + 1];
    }
    free(start);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Join helpers                                                              */
/* ------------------------------------------------------------------------ */

/* Prefix lengths are clamped to the record length.  For every non-empty record
 * max_prefix and mid_prefix are already <= |x|; an empty record would make the
 * reference read one token past its end (max_prefix(0) = 1, similarity.cuh:44-47,
 * checking.cu:126), so here it simply has no prefix and never joins. */
static uint32_t prefix_clamp(uint32_t p, uint32_t len) { return p < len ? p : len; }

/* Inverted index over mid-prefix tokens (index.cu:169-211 with the entry
 * kernel :276-296).  Lists hold (set, pos) grouped by term, sets ascending. */
typedef struct { int64_t *off; int32_t *sid; int32_t *spos; } orc_index;

static uint32_t index_prefix(uint32_t l, float t, int use_max)
{
    return prefix_clamp(use_max ? orc_max_prefix(l, t) : orc_mid_prefix(l, t), l);
}

static void build_index_sel(int64_t n, const int32_t *len, const int64_t *pos,
                            const int32_t *tok, int32_t n_terms, float t, int use_max,
                            orc_index *ix, int64_t *e_idx)
{
    ix->off = (int64_t *)calloc((size_t)n_terms + 1, sizeof(int64_t));
    int64_t E = 0;
    for (int64_t s = 0; s < n; s++) {
        uint32_t mp = index_prefix((uint32_t)len[s], t, use_max);   /* data.cu:241-257 */
        for (uint32_t j = 0; j < mp; j++) ix->off[tok[pos[s] + j] + 1]++;
        E += mp;
    }
    for (int32_t k = 0; k < n_terms; k++) ix->off[k + 1] += ix->off[k];
    ix->sid = (int32_t *)malloc((size_t)(E > 0 ? E : 1) * sizeof(int32_t));
    ix->spos = (int32_t *)malloc((size_t)(E > 0 ? E : 1) * sizeof(int32_t));
    int64_t *fill = (int64_t *)malloc((size_t)(n_terms > 0 ? n_terms : 1) * sizeof(int64_t));
    memcpy(fill, ix->off, (size_t)n_terms * sizeof(int64_t));
    for (int64_t s = 0; s < n; s++) {
        uint32_t mp = index_prefix((uint32_t)len[s], t, use_max);
        for (uint32_t j = 0; j < mp; j++) {
            int64_t at = fill[tok[pos[s] + j]]++;
            ix->sid[at] = (int32_t)s; ix->spos[at] = (int32_t)j;
        }
    }
    free(fill);
    *e_idx = E;
}

static void build_index(int64_t n, const int32_t *len, const int64_t *pos,
                        const int32_t *tok, int32_t n_terms, float t, orc_index *ix,
                        int64_t *e_idx)
{
    build_index_sel(n, len, pos, tok, n_terms, t, 0, ix, e_idx);
}

static void free_index(orc_index *ix) { free(ix->off); free(ix->sid); free(ix->spos); }

/* The 4-way size filter, filtering.cu:88-92 and :111-120 (floor of the float
 * bound, then an unsigned compare). */
static int size_ok(uint32_t lq, uint32_t ls, float t)
{
    uint32_t qmin = (uint32_t)floorf(orc_min_size(lq, t));
    uint32_t qmax = (uint32_t)floorf(orc_max_size(lq, t));
    uint32_t smin = (uint32_t)floorf(orc_min_size(ls, t));
    uint32_t smax = (uint32_t)floorf(orc_max_size(ls, t));
    return !(ls < qmin || ls > qmax || lq < smin || lq > smax);
}

/* Verification exactly as checking_kernel_block (checking.cu:93-175).
 * Returns the stored score (overlap, or 0 when below floor(bound));
 * *reads counts the token loads (2 boundary + 2 per loop iteration). */
static uint32_t rq_verify(const int32_t *Q, uint32_t lq, const int32_t *S, uint32_t ls,
                          float t, uint32_t partial, int64_t *reads)
{
    int minoverlap = (int)floorf(orc_min_overlap(lq, ls, t));      /* :106 */
    uint32_t p1 = 0, p2 = 0;
    uint32_t score = (uint16_t)partial;                           /* :120 */
    uint32_t qp = orc_max_prefix(lq, t), sp = orc_mid_prefix(ls, t);
    uint32_t tk1 = (uint32_t)Q[qp - 1], tk2 = (uint32_t)S[sp - 1];  /* :126-127 */
    int64_t rd = 2;
    if (tk1 <= tk2) p1 = qp; else p2 = sp;                        /* :129-130 */
    while (p1 < lq && p2 < ls) {                                  /* :133 */
        uint32_t a = (uint32_t)Q[p1], b = (uint32_t)S[p2];
        rd += 2;
        if ((p1 == lq - 1 && a < b) || (p2 == ls - 1 && b < a)) break;   /* :138-140 */
        if (a == b) { score = (uint16_t)(score + 1); p1++; p2++; }
        else {
            uint32_t rem = (a < b) ? lq - p1 : ls - p2;
            if ((rem + score) < (uint32_t)minoverlap) { score = 0; break; }  /* :156-160 */
            if (a < b) p1++; else p2++;
        }
    }
    if (reads) *reads += rd;
    if ((int)score < minoverlap) return 0;                         /* :167-173 */
    return score;
}

/* |Q intersect S| by a plain merge (the closed form of the verification). */
static uint32_t overlap_full(const int32_t *Q, uint32_t lq, const int32_t *S, uint32_t ls)
{
    uint32_t i = 0, j = 0, c = 0;
    while (i < lq && j < ls) {
        if (Q[i] == S[j]) { c++; i++; j++; }
        else if (Q[i] < S[j]) i++; else j++;
    }
    return c;
}

/* token reads of a merge restricted to tokens > tau = min(Q[qp-1], S[sp-1]):
 * every common token <= tau is already counted by the prefix overlap. */
static int64_t suffix_reads(const int32_t *Q, uint32_t lq, const int32_t *S, uint32_t ls,
                            uint32_t qp, uint32_t sp)
{
    int32_t tau = Q[qp - 1] < S[sp - 1] ? Q[qp - 1] : S[sp - 1];
    uint32_t a = 0, b = 0;
    while (a < lq && Q[a] <= tau) a++;
    while (b < ls && S[b] <= tau) b++;
    return (int64_t)(lq - a) + (int64_t)(ls - b);
}

/* growable triple buffer */
typedef struct { int32_t *v; int64_t n, cap; } tri_buf;
static void tri_push(tri_buf *b, int32_t q, int32_t s, int32_t ov)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->v = (int32_t *)realloc(b->v, (size_t)b->cap * 3 * sizeof(int32_t));
    }
    b->v[3 * b->n] = q; b->v[3 * b->n + 1] = s; b->v[3 * b->n + 2] = ov; b->n++;
}
static int cmp_tri(const void *a, const void *b)
{
    const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Closed-form join, multithreaded over query chunks                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n; const int32_t *len; const int64_t *pos; const int32_t *tok; float t;
    const orc_index *ix;
    const int64_t *queries; int64_t nq;   /* a query subset (NULL: every record) */
    int checks;                           /* 0: the reference's merge only, once per candidate */
    atomic_llong next;
} cf_shared;

typedef struct {
    cf_shared *sh;
    tri_buf out;
    orc_counters c;
} cf_worker;

static void *cf_run(void *arg)
{
    cf_worker *w = (cf_worker *)arg;
    cf_shared *sh = w->sh;
    const int64_t n = sh->n;
    const float t = sh->t;
    int32_t *cnt = (int32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *lastj = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int32_t *lasts = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int32_t *touched = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    const int64_t nq = sh->queries ? sh->nq : n;
    /* queries per claim: 32 over a whole collection, 1 over a subset (a sampled
     * subset of a few hundred queries of very different cost would otherwise leave
     * most threads idle) */
    const int64_t CH = sh->queries ? 1 : 32;
    for (;;) {
        int64_t q0 = atomic_fetch_add(&sh->next, CH);
        if (q0 >= nq) break;
        int64_t q1 = q0 + CH < nq ? q0 + CH : nq;
        for (int64_t qk = q0; qk < q1; qk++) {
            const int64_t q = sh->queries ? sh->queries[qk] : qk;
            const uint32_t lq = (uint32_t)sh->len[q];
            const int32_t *Q = sh->tok + sh->pos[q];
            const uint32_t qp = prefix_clamp(orc_max_prefix(lq, t), lq);
            int64_t nt = 0;
            w->c.p_q += qp;
            /* filtering.cu:94-139: probe-prefix token j, every posting (s, spos) */
            for (uint32_t j = 0; j < qp; j++) {
                int32_t tk = Q[j];
                int64_t a = sh->ix->off[tk], b = sh->ix->off[tk + 1];
                /* lists are ascending in s: skip s <= q (filtering.cu:107) */
                int64_t lo = a, hi = b;
                while (lo < hi) { int64_t mid = (lo + hi) / 2; if (sh->ix->sid[mid] <= q) lo = mid + 1; else hi = mid; }
                for (int64_t e = lo; e < b; e++) {
                    int32_t s = sh->ix->sid[e];
                    uint32_t ls = (uint32_t)sh->len[s];
                    if (!size_ok(lq, ls, t)) continue;
                    w->c.e_scan++;
                    if (cnt[s] == 0) touched[nt++] = s;
                    cnt[s]++; lastj[s] = (int32_t)j; lasts[s] = sh->ix->spos[e];
                }
            }
            /* rules 3-6 of the closed form */
            for (int64_t k = 0; k < nt; k++) {
                int32_t s = touched[k];
                uint32_t m = (uint32_t)cnt[s];
                cnt[s] = 0;
                uint32_t ls = (uint32_t)sh->len[s];
                const int32_t *S = sh->tok + sh->pos[s];
                float bound = orc_min_overlap(lq, ls, t);
                uint32_t rq = lq - (uint32_t)lastj[s], rs = ls - (uint32_t)lasts[s];
                uint32_t minrem = rq < rs ? rq : rs;
                if ((float)(m + minrem) < bound) continue;        /* positional filter */
                w->c.candidates++;
                if (!sh->checks) {   /* checking.cu:93-175 alone: its score is the overlap or 0 */
                    uint32_t sc = rq_verify(Q, lq, S, ls, t, m, &w->c.m_reads);
                    if (sc > 0) {
                        w->c.intersections++;
                        if (orc_similarity(lq, ls, sc) >= t) {             /* block.cu:189-192 */
                            tri_push(&w->out, (int32_t)q, s, (int32_t)sc);
                            w->c.results++;
                        }
                    }
                    continue;
                }
                uint32_t ov = overlap_full(Q, lq, S, ls);
                uint32_t rqs = rq_verify(Q, lq, S, ls, t, m, &w->c.m_reads);
                w->c.v_reads += suffix_reads(Q, lq, S, ls, qp, orc_mid_prefix(ls, t));
                uint32_t want = ((int)ov >= (int)floorf(bound)) ? ov : 0;
                if (rqs != want) {
                    fprintf(stderr, "orc_selfjoin: verification mismatch at (%lld,%d)\n", (long long)q, s);
                    abort();
                }
                if (ov > 0 && (int)ov >= (int)floorf(bound)) {
                    w->c.intersections++;
                    if (orc_similarity(lq, ls, ov) >= t) {                 /* block.cu:189-192 */
                        tri_push(&w->out, (int32_t)q, s, (int32_t)ov);
                        w->c.results++;
                    }
                }
            }
        }
    }
    free(cnt); free(lastj); free(lasts); free(touched);
    return NULL;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int64_t selfjoin_closed(int64_t n_sets, const int32_t *len, const int64_t *pos,
                               const int32_t *tok, int32_t n_terms, float t, int nthreads,
                               const int64_t *queries, int64_t nq, int checks, int32_t **out, orc_counters *ctr,
                               double *secs)
{
    if (nthreads < 1) nthreads = 1;
    orc_index ix;
    orc_counters total;
    memset(&total, 0, sizeof(total));
    const double t0 = now_s();
    build_index(n_sets, len, pos, tok, n_terms, t, &ix, &total.e_idx);
    const double t1 = now_s();
    cf_shared sh;
    sh.n = n_sets; sh.len = len; sh.pos = pos; sh.tok = tok; sh.t = t; sh.ix = &ix;
    sh.queries = queries; sh.nq = nq; sh.checks = checks;
    atomic_init(&sh.next, 0);
    cf_worker *ws = (cf_worker *)calloc((size_t)nthreads, sizeof(cf_worker));
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int i = 0; i < nthreads; i++) { ws[i].sh = &sh; pthread_create(&th[i], NULL, cf_run, &ws[i]); }
    int64_t R = 0;
    for (int i = 0; i < nthreads; i++) { pthread_join(th[i], NULL); R += ws[i].out.n; }
    int32_t *res = (int32_t *)malloc((size_t)(R > 0 ? R : 1) * 3 * sizeof(int32_t));
    int64_t o = 0;
    for (int i = 0; i < nthreads; i++) {
        if (ws[i].out.n) memcpy(res + 3 * o, ws[i].out.v, (size_t)ws[i].out.n * 3 * sizeof(int32_t));
        o += ws[i].out.n;
        free(ws[i].out.v);
        total.p_q += ws[i].c.p_q; total.e_scan += ws[i].c.e_scan; total.m_reads += ws[i].c.m_reads;
        total.v_reads += ws[i].c.v_reads; total.candidates += ws[i].c.candidates;
        total.intersections += ws[i].c.intersections; total.results += ws[i].c.results;
    }
    const double t2 = now_s();
    qsort(res, (size_t)R, 3 * sizeof(int32_t), cmp_tri);
    free(ws); free(th); free_index(&ix);
    *out = res;
    if (ctr) *ctr = total;
    if (secs) { secs[0] = t1 - t0; secs[1] = t2 - t1; }
    return R;
}

int64_t orc_selfjoin(int64_t n_sets, const int32_t *len, const int64_t *pos,
                     const int32_t *tok, int32_t n_terms, float t, int nthreads,
                     int32_t **out, orc_counters *ctr)
{
    return selfjoin_closed(n_sets, len, pos, tok, n_terms, t, nthreads, NULL, 0, 1, out, ctr, NULL);
}

int64_t orc_selfjoin_ex(int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                        int32_t n_terms, float t, int nthreads, const int64_t *queries, int64_t n_queries,
                        int checks, int32_t **out, orc_counters *ctr, double *secs)
{
    for (int64_t k = 0; queries && k < n_queries; k++)
        if (queries[k] < 0 || queries[k] >= n_sets) return -1;
    return selfjoin_closed(n_sets, len, pos, tok, n_terms, t, nthreads, queries, queries ? n_queries : 0, checks,
                           out, ctr, secs);
}

int64_t orc_selfjoin_queries(int64_t n_sets, const int32_t *len, const int64_t *pos,
                             const int32_t *tok, int32_t n_terms, float t, const int64_t *queries,
                             int64_t n_queries, int nthreads, int32_t **out, orc_counters *ctr)
{
    for (int64_t k = 0; k < n_queries; k++)
        if (queries[k] < 0 || queries[k] >= n_sets) return -1;
    return selfjoin_closed(n_sets, len, pos, tok, n_terms, t, nthreads, queries, n_queries, 1, out, ctr, NULL);
}

/* ------------------------------------------------------------------------ */
/* Literal restatement of the reference tile loop (block.cu:61-213)          */
/* ------------------------------------------------------------------------ */
int64_t orc_selfjoin_literal(int64_t n_sets, const int32_t *len, const int64_t *pos,
                             const int32_t *tok, int32_t n_terms, float t,
                             int64_t block_size, int32_t **out, orc_counters *ctr)
{
    orc_counters c;
    memset(&c, 0, sizeof(c));
    if (block_size < 1) return -1;
    if (block_size > n_sets) block_size = n_sets;          /* block.cu:50-51 */
    orc_index ix;
    build_index(n_sets, len, pos, tok, n_terms, t, &ix, &c.e_idx);
    const int64_t B = block_size;
    int64_t n_blocks = n_sets / B + (n_sets % B ? 1 : 0);  /* block.cu:77-81 */
    int64_t last_block = n_sets % B ? n_sets % B : B;
    int16_t *scores = (int16_t *)malloc((size_t)(B * B > 0 ? B * B : 1) * sizeof(int16_t));
    int64_t *cand = (int64_t *)malloc((size_t)(B * B > 0 ? B * B : 1) * sizeof(int64_t));
    tri_buf res = {0};
    for (int64_t i = 0; i < n_blocks; i++) {
        int64_t i_off = i * B;
        for (int64_t j = 0; j <= i; j++) {
            int64_t q_off = j * B;
            int64_t bsize = (j == n_blocks - 1) ? last_block : B;
            /* block.cu:121-130: tile skip */
            int q_size = len[q_off + bsize - 1];
            int s_maxsize = (int)floorf(orc_max_size((uint32_t)len[i_off], t));
            if (q_size > s_maxsize) continue;
            memset(scores, 0, (size_t)(B * B) * sizeof(int16_t));   /* filtering.cu:23 */
            /* filtering_kernel_block, one query at a time (filtering.cu:83-140) */
            for (int64_t qq = 0; qq < bsize; qq++) {
                int64_t query = qq + q_off;
                uint32_t lq = (uint32_t)len[query];
                const int32_t *Q = tok + pos[query];
                uint32_t qp = prefix_clamp(orc_max_prefix(lq, t), lq);
                for (uint32_t jj = 0; jj < qp; jj++) {
                    int32_t tk = Q[jj];
                    for (int64_t e = ix.off[tk]; e < ix.off[tk + 1]; e++) {
                        int64_t source = ix.sid[e];
                        uint32_t stpos = (uint32_t)ix.spos[e];
                        if (query < source && source >= i_off && source < i_off + B) {
                            uint32_t ls = (uint32_t)len[source];
                            if (!size_ok(lq, ls, t)) continue;
                            c.e_scan++;
                            int64_t bucket = (query - q_off) * B + (source - i_off);
                            int16_t score = scores[bucket];
                            if (score >= 0) {
                                uint32_t qr = lq - jj, sr = ls - stpos;
                                uint32_t minrem = qr < sr ? qr : sr;
                                float mo = orc_min_overlap(lq, ls, t);
                                if ((float)((uint32_t)(score + 1) + minrem) < mo) scores[bucket] = -1;
                                else scores[bucket] = (int16_t)(score + 1);
                            }
                        }
                    }
                }
            }
            /* filter_k (compaction.cu:8-17): buckets with score > 0 */
            int64_t nc = 0;
            for (int64_t b = 0; b < B * B; b++) if (scores[b] > 0) cand[nc++] = b;
            c.candidates += nc;
            /* checking_kernel_block + host render (block.cu:183-197) */
            for (int64_t k = 0; k < nc; k++) {
                int64_t b = cand[k];
                int64_t query = b / B + q_off, source = b % B + i_off;
                uint32_t lq = (uint32_t)len[query], ls = (uint32_t)len[source];
                uint32_t sc = rq_verify(tok + pos[query], lq, tok + pos[source], ls, t,
                                        (uint32_t)scores[b], &c.m_reads);
                if (sc > 0) {
                    c.intersections++;
                    if (orc_similarity(lq, ls, sc) >= t) {
                        tri_push(&res, (int32_t)query, (int32_t)source, (int32_t)sc);
                        c.results++;
                    }
                }
            }
        }
    }
    for (int64_t q = 0; q < n_sets; q++) c.p_q += prefix_clamp(orc_max_prefix((uint32_t)len[q], t), (uint32_t)len[q]);
    free(scores); free(cand); free_index(&ix);
    if (res.n > 1) qsort(res.v, (size_t)res.n, 3 * sizeof(int32_t), cmp_tri);   /* res.v is null when empty */
    *out = res.v ? res.v : (int32_t *)malloc(sizeof(int32_t));
    if (ctr) *ctr = c;
    return res.n;
}

/* ------------------------------------------------------------------------ */
/* Brute force: all q<s with float32 Jaccard >= t                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n; const int32_t *len; const int64_t *pos; const int32_t *tok; float t;
    atomic_llong next;
} bf_shared;
typedef struct { bf_shared *sh; tri_buf out; } bf_worker;

static void *bf_run(void *arg)
{
    bf_worker *w = (bf_worker *)arg;
    bf_shared *sh = w->sh;
    for (;;) {
        int64_t q = atomic_fetch_add(&sh->next, 1);
        if (q >= sh->n) break;
        for (int64_t s = q + 1; s < sh->n; s++) {
            uint32_t lq = (uint32_t)sh->len[q], ls = (uint32_t)sh->len[s];
            uint32_t ov = overlap_full(sh->tok + sh->pos[q], lq, sh->tok + sh->pos[s], ls);
            if (ov > 0 && orc_similarity(lq, ls, ov) >= sh->t) tri_push(&w->out, (int32_t)q, (int32_t)s, (int32_t)ov);
        }
    }
    return NULL;
}

int64_t orc_selfjoin_brute(int64_t n_sets, const int32_t *len, const int64_t *pos,
                           const int32_t *tok, float t, int nthreads, int32_t **out)
{
    if (nthreads < 1) nthreads = 1;
    bf_shared sh;
    sh.n = n_sets; sh.len = len; sh.pos = pos; sh.tok = tok; sh.t = t;
    atomic_init(&sh.next, 0);
    bf_worker *ws = (bf_worker *)calloc((size_t)nthreads, sizeof(bf_worker));
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int i = 0; i < nthreads; i++) { ws[i].sh = &sh; pthread_create(&th[i], NULL, bf_run, &ws[i]); }
    int64_t R = 0;
    for (int i = 0; i < nthreads; i++) { pthread_join(th[i], NULL); R += ws[i].out.n; }
    int32_t *res = (int32_t *)malloc((size_t)(R > 0 ? R : 1) * 3 * sizeof(int32_t));
    int64_t o = 0;
    for (int i = 0; i < nthreads; i++) {
        if (ws[i].out.n) memcpy(res + 3 * o, ws[i].out.v, (size_t)ws[i].out.n * 3 * sizeof(int32_t));
        o += ws[i].out.n; free(ws[i].out.v);
    }
    qsort(res, (size_t)R, 3 * sizeof(int32_t), cmp_tri);
    free(ws); free(th);
    *out = res;
    return R;
}

/* ------------------------------------------------------------------------ */
/* R-S join (probe records against an indexed collection)                    */
/* ------------------------------------------------------------------------ */
/* The reference is a self-join only; its R-S mode exists as a usage string
 * (src/preprocess_cli.cpp:26).  The R-S rules are the self-join's closed form
 * (filtering.cu:76-141, checking.cu:93-175, block.cu:183-197) with the longer record
 * of each pair as the query, the probe on ties: a probe r and an indexed s with
 * |S| <= |R| use qp = max_prefix(|R|, t), sp = mid_prefix(|S|, t); with |S| > |R|,
 * qp = mid_prefix(|R|, t), sp = max_prefix(|S|, t).  (r, s) is a result iff
 * size_ok(|R|, |S|), m = |R[0,qp) n S[0,sp)| > 0, the positional filter passes at
 * the last prefix match, |R n S| >= floor(min_overlap) and the float32 Jaccard is
 * >= t.  The index holds max_prefix(|S|, t_index) tokens (t_index <= t); e_scan
 * counts the postings of the probe's prefix tokens (qp of each half) inside the
 * size band before postings past sp are skipped. */
typedef struct {
    int64_t np; const int32_t *plen; const int64_t *ppos; const int32_t *ptok;
    int64_t n; const int32_t *len; const int64_t *pos; const int32_t *tok; int32_t n_terms;
    float t;
    const orc_index *ix;
    atomic_llong next;
} rs_shared;

typedef struct { rs_shared *sh; tri_buf out; orc_counters c; } rs_worker;

static void *rs_run(void *arg)
{
    rs_worker *w = (rs_worker *)arg;
    rs_shared *sh = w->sh;
    const float t = sh->t;
    const int64_t n = sh->n;
    int32_t *cnt = (int32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *lastj = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int32_t *lasts = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int32_t *touched = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    for (;;) {
        int64_t r = atomic_fetch_add(&sh->next, 1);
        if (r >= sh->np) break;
        const uint32_t lr = (uint32_t)sh->plen[r];
        const int32_t *R = sh->ptok + sh->ppos[r];
        const uint32_t qmax = prefix_clamp(orc_max_prefix(lr, t), lr);
        const uint32_t qmid = prefix_clamp(orc_mid_prefix(lr, t), lr);
        int64_t nt = 0;
        w->c.p_q += qmax;
        for (uint32_t j = 0; j < qmax; j++) {
            int32_t tk = R[j];
            if (tk >= sh->n_terms) continue;            /* a token the index never saw */
            for (int64_t e = sh->ix->off[tk]; e < sh->ix->off[tk + 1]; e++) {
                int32_t s = sh->ix->sid[e];
                uint32_t ls = (uint32_t)sh->len[s];
                if (!size_ok(lr, ls, t)) continue;
                const int longer = ls > lr;              /* s is the query of the pair */
                if (longer && j >= qmid) continue;
                w->c.e_scan++;
                uint32_t sp = prefix_clamp(longer ? orc_max_prefix(ls, t) : orc_mid_prefix(ls, t), ls);
                if ((uint32_t)sh->ix->spos[e] >= sp) continue;
                if (cnt[s] == 0) touched[nt++] = s;
                cnt[s]++; lastj[s] = (int32_t)j; lasts[s] = sh->ix->spos[e];
            }
        }
        for (int64_t k = 0; k < nt; k++) {
            int32_t s = touched[k];
            uint32_t m = (uint32_t)cnt[s];
            cnt[s] = 0;
            uint32_t ls = (uint32_t)sh->len[s];
            float bound = orc_min_overlap(lr, ls, t);
            uint32_t rq = lr - (uint32_t)lastj[s], rs = ls - (uint32_t)lasts[s];
            uint32_t minrem = rq < rs ? rq : rs;
            if ((float)(m + minrem) < bound) continue;
            w->c.candidates++;
            uint32_t ov = overlap_full(R, lr, sh->tok + sh->pos[s], ls);
            if (ov > 0 && (int)ov >= (int)floorf(bound)) {
                w->c.intersections++;
                if (orc_similarity(lr, ls, ov) >= t) {
                    tri_push(&w->out, (int32_t)r, s, (int32_t)ov);
                    w->c.results++;
                }
            }
        }
    }
    free(cnt); free(lastj); free(lasts); free(touched);
    return NULL;
}

static int64_t gather(tri_buf *bufs, size_t stride, int nthreads, int32_t **out)
{
    int64_t R = 0;
    for (int i = 0; i < nthreads; i++) R += ((tri_buf *)((char *)bufs + (size_t)i * stride))->n;
    int32_t *res = (int32_t *)malloc((size_t)(R > 0 ? R : 1) * 3 * sizeof(int32_t));
    int64_t o = 0;
    for (int i = 0; i < nthreads; i++) {
        tri_buf *b = (tri_buf *)((char *)bufs + (size_t)i * stride);
        if (b->n) memcpy(res + 3 * o, b->v, (size_t)b->n * 3 * sizeof(int32_t));
        o += b->n;
        free(b->v);
    }
    qsort(res, (size_t)R, 3 * sizeof(int32_t), cmp_tri);
    *out = res;
    return R;
}

int64_t orc_rsjoin(int64_t n_probes, const int32_t *plen, const int64_t *ppos, const int32_t *ptok,
                   int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                   int32_t n_terms, float t, float t_index, int nthreads, int32_t **out,
                   orc_counters *ctr)
{
    if (nthreads < 1) nthreads = 1;
    if (!(t_index <= t)) return -1;
    orc_index ix;
    orc_counters total;
    memset(&total, 0, sizeof(total));
    build_index_sel(n_sets, len, pos, tok, n_terms, t_index, 1, &ix, &total.e_idx);
    rs_shared sh;
    sh.np = n_probes; sh.plen = plen; sh.ppos = ppos; sh.ptok = ptok;
    sh.n = n_sets; sh.len = len; sh.pos = pos; sh.tok = tok; sh.n_terms = n_terms; sh.t = t; sh.ix = &ix;
    atomic_init(&sh.next, 0);
    rs_worker *ws = (rs_worker *)calloc((size_t)nthreads, sizeof(rs_worker));
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int i = 0; i < nthreads; i++) { ws[i].sh = &sh; pthread_create(&th[i], NULL, rs_run, &ws[i]); }
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    for (int i = 0; i < nthreads; i++) {
        total.p_q += ws[i].c.p_q; total.e_scan += ws[i].c.e_scan; total.candidates += ws[i].c.candidates;
        total.intersections += ws[i].c.intersections; total.results += ws[i].c.results;
    }
    int64_t R = gather(&ws[0].out, sizeof(rs_worker), nthreads, out);
    free(ws); free(th); free_index(&ix);
    if (ctr) *ctr = total;
    return R;
}

/* every (r, s) with float32 Jaccard >= t and a common token (a superset of orc_rsjoin) */
int64_t orc_rsjoin_brute(int64_t n_probes, const int32_t *plen, const int64_t *ppos, const int32_t *ptok,
                         int64_t n_sets, const int32_t *len, const int64_t *pos, const int32_t *tok,
                         float t, int32_t **out)
{
    tri_buf b = {0};
    for (int64_t r = 0; r < n_probes; r++)
        for (int64_t s = 0; s < n_sets; s++) {
            uint32_t lr = (uint32_t)plen[r], ls = (uint32_t)len[s];
            uint32_t ov = overlap_full(ptok + ppos[r], lr, tok + pos[s], ls);
            if (ov > 0 && orc_similarity(lr, ls, ov) >= t) tri_push(&b, (int32_t)r, (int32_t)s, (int32_t)ov);
        }
    return gather(&b, sizeof(tri_buf), 1, out);
}

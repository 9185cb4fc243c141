/* oracle_check.c -- the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
 * tokenize a text, load the .bin stream it makes, and run every join of the oracle on it
 * at several thresholds: the closed form (with and without its self-checks, threaded and
 * on a query subset), the reference's tile loop (block sizes 1, 7 and n), brute force,
 * and the R-S join of a probe half against the collection; the results must agree as
 * they do in the Python tests (closed form = tile loop, closed form within brute force).
 *
 * usage: oracle_check <text file> <q> <t> [<t> ...]
 * Exit status 0 when every check held (a sanitizer report aborts the run on its own). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fgss_oracle.h"

static int g_fail = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
            g_fail++;                                                             \
        }                                                                         \
    } while (0)

static char *read_text(const char *path, int64_t *n)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 16, len = 0;
    char *b = malloc(cap);
    size_t k;
    while (b && (k = fread(b + len, 1, cap - len, f)) > 0) {
        len += k;
        if (len == cap) {
            char *nb = realloc(b, cap * 2);
            if (!nb) { free(b); b = NULL; break; }
            b = nb;
            cap *= 2;
        }
    }
    fclose(f);
    *n = (int64_t)len;
    return b;
}

/* triples a (na) within triples b (nb), both sorted by (q, s) */
static int subset(const int32_t *a, int64_t na, const int32_t *b, int64_t nb)
{
    int64_t j = 0;
    for (int64_t i = 0; i < na; i++) {
        while (j < nb && (b[3 * j] < a[3 * i] || (b[3 * j] == a[3 * i] && b[3 * j + 1] < a[3 * i + 1]))) j++;
        if (j == nb || b[3 * j] != a[3 * i] || b[3 * j + 1] != a[3 * i + 1] || b[3 * j + 2] != a[3 * i + 2]) return 0;
    }
    return 1;
}

static int same_ctr(const orc_counters *a, const orc_counters *b)
{
    return a->e_idx == b->e_idx && a->p_q == b->p_q && a->candidates == b->candidates &&
           a->intersections == b->intersections && a->results == b->results;
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <text file> <q> <t> [<t> ...]\n", argv[0]);
        return 2;
    }
    int64_t nbytes = 0;
    char *text = read_text(argv[1], &nbytes);
    if (!text) {
        fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    const int q = atoi(argv[2]);
    int32_t *bin = NULL, *loc = NULL;
    int64_t nbin = 0, nloc = 0;
    CHECK(orc_tokenize(text, nbytes, q, &bin, &nbin, &loc, &nloc) == 0);
    int64_t n = 0, ntok = 0;
    int32_t nterms = 0;
    CHECK(orc_bin_count(bin, nbin, &n, &ntok, &nterms) == 0);
    CHECK(nloc == ntok);
    int32_t *id = malloc(sizeof(int32_t) * (size_t)(n + 1)), *len = malloc(sizeof(int32_t) * (size_t)(n + 1));
    int64_t *pos = malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t *tok = malloc(sizeof(int32_t) * (size_t)(ntok + 1));
    CHECK(orc_bin_load(bin, nbin, id, len, pos, tok) == 0);
    int64_t total = 0;
    for (int k = 3; k < argc; k++) {
        const float t = orc_parse_threshold(argv[k]);
        orc_counters c0, c1, c2, c3;
        int32_t *r0 = NULL, *r1 = NULL, *r2 = NULL, *rb = NULL;
        memset(&c0, 0, sizeof c0);
        const int64_t n0 = orc_selfjoin(n, len, pos, tok, nterms, t, 4, &r0, &c0);
        const int64_t n1 = orc_selfjoin_ex(n, len, pos, tok, nterms, t, 3, NULL, 0, 0, &r1, &c1, NULL);
        CHECK(n0 >= 0 && n1 == n0 && same_ctr(&c0, &c1));
        CHECK(n0 == 0 || memcmp(r0, r1, sizeof(int32_t) * 3 * (size_t)n0) == 0);
        const int64_t blocks[3] = {1, 7, n > 0 ? n : 1};
        for (int b = 0; b < 3; b++) {
            const int64_t n2 = orc_selfjoin_literal(n, len, pos, tok, nterms, t, blocks[b], &r2, &c2);
            CHECK(n2 == n0 && same_ctr(&c0, &c2));
            CHECK(n0 == 0 || (n2 == n0 && memcmp(r0, r2, sizeof(int32_t) * 3 * (size_t)n0) == 0));
            orc_free(r2);
            r2 = NULL;
        }
        const int64_t nbr = orc_selfjoin_brute(n, len, pos, tok, t, 2, &rb);
        CHECK(nbr >= n0 && subset(r0, n0, rb, nbr));
        /* the odd queries alone: their share of the closed form */
        int64_t nq = n / 2;
        int64_t *qs = malloc(sizeof(int64_t) * (size_t)(nq + 1));
        for (int64_t i = 0; i < nq; i++) qs[i] = 2 * i + 1;
        int32_t *r3 = NULL;
        const int64_t n3 = orc_selfjoin_queries(n, len, pos, tok, nterms, t, qs, nq, 2, &r3, &c3);
        CHECK(n3 >= 0 && subset(r3, n3, r0, n0));
        int64_t odd = 0;
        for (int64_t i = 0; i < n0; i++) odd += r0[3 * i] & 1;
        CHECK(n3 == odd);
        /* R-S: the first half of the records (any order) probing the whole collection */
        const int64_t np = n / 2;
        int32_t *rr = NULL, *rrb = NULL;
        orc_counters cr;
        const int64_t nr = orc_rsjoin(np, len, pos, tok, n, len, pos, tok, nterms, t, t, 2, &rr, &cr);
        const int64_t nrb = orc_rsjoin_brute(np, len, pos, tok, n, len, pos, tok, t, &rrb);
        CHECK(nr >= 0 && nrb >= nr && subset(rr, nr, rrb, nrb));
        total += n0 + n3 + nr;
        orc_free(r0);
        orc_free(r1);
        orc_free(rb);
        orc_free(r3);
        orc_free(rr);
        orc_free(rrb);
        free(qs);
    }
    printf("oracle_check: %lld records, %lld triples, %d failed checks\n", (long long)n, (long long)total, g_fail);
    free(id);
    free(len);
    free(pos);
    free(tok);
    orc_free(bin);
    orc_free(loc);
    free(text);
    return g_fail ? 1 : 0;
}

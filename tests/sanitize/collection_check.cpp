// collection_check.cpp -- the host collection code under AddressSanitizer and
// UndefinedBehaviorSanitizer: the .bin reader on the reference's fixture, on hand-made
// malformed streams and on seeded mutations of the fixture; the text front end at
// q = 1..12 on the fixture text, edge texts and generated text; probes tokenized against
// a dictionary; the .bin/.loc writer and a reload of what it wrote.
//
// usage: collection_check <fixture .bin> <fixture text> <scratch dir>
// Exit status 0 when every check held (a sanitizer report aborts the run on its own).
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "fgssjoin.h"

static int g_fail = 0;
static long g_loads = 0;

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
            g_fail++;                                                            \
        }                                                                        \
    } while (0)

static std::vector<char> read_bytes(const char *path)
{
    std::vector<char> b;
    FILE *f = fopen(path, "rb");
    if (!f) return b;
    char chunk[4096];
    size_t k;
    while ((k = fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + k);
    fclose(f);
    return b;
}

// Walk every array of a loaded collection once (a wrong length or offset would read
// outside them) and check the invariants the join relies on.
static void walk(const fgss_collection *c)
{
    fgss_collection_info info;
    CHECK(fgss_collection_get_info(c, &info) == FGSS_OK);
    const int32_t *id = nullptr, *len = nullptr, *tok = nullptr;
    const int64_t *pos = nullptr;
    CHECK(fgss_collection_arrays(c, &id, &len, &pos, &tok) == FGSS_OK);
    int64_t sum = 0;
    int32_t mx = 0;
    uint64_t h = 0;
    for (int64_t i = 0; i < info.n_sets; i++) {
        CHECK(pos[i] == sum);
        CHECK(len[i] >= 0);
        h = h * 31u + (uint32_t)id[i];
        for (int32_t k = 0; k < len[i]; k++) {
            const int32_t v = tok[pos[i] + k];
            CHECK(v < info.n_terms);
            h = h * 131u + (uint32_t)v;
        }
        sum += len[i];
        if (len[i] > mx) mx = len[i];
    }
    CHECK(sum == info.n_tokens);
    CHECK(mx == info.max_len);
    if (h == 42) fprintf(stderr, "%llu\n", (unsigned long long)h);   // keeps the reads
}

// Load an int32 stream; on success walk it.  Returns the status.
static int load(const std::vector<int32_t> &v)
{
    g_loads++;
    fgss_collection *c = nullptr;
    const int rc = fgss_collection_from_bin_buffer(v.data(), (int64_t)v.size(), &c);
    if (rc == FGSS_OK) {
        CHECK(c != nullptr);
        walk(c);
        fgss_collection_free(c);
    } else {
        CHECK(c == nullptr);
        CHECK(fgss_last_error()[0] != '\0');
    }
    return rc;
}

static uint64_t g_rng = 0x9E3779B97F4A7C15ull;
static uint64_t next_rand()
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return g_rng;
}

static void check_bin(const std::vector<int32_t> &fix)
{
    CHECK(load(fix) == FGSS_OK);
    {
        fgss_collection *c = nullptr;
        CHECK(fgss_collection_from_bin_buffer(fix.data(), (int64_t)fix.size(), &c) == FGSS_OK);
        fgss_collection_info info;
        fgss_collection_get_info(c, &info);
        CHECK(info.n_sets == 5 && info.n_tokens == 88 && info.n_terms == 42);
        fgss_collection_free(c);
    }
    // hand-made streams: empty, truncated headers and bodies, negative and huge lengths,
    // token ids at the int32 limits, unordered and repeated tokens
    CHECK(load({}) == FGSS_OK);
    CHECK(load({7}) == FGSS_ERR_FORMAT);
    CHECK(load({7, -1}) == FGSS_ERR_FORMAT);
    CHECK(load({7, INT_MIN}) == FGSS_ERR_FORMAT);
    CHECK(load({7, INT_MAX, 1, 2}) == FGSS_ERR_FORMAT);
    CHECK(load({7, 3, 1, 2}) == FGSS_ERR_FORMAT);
    CHECK(load({7, 0}) == FGSS_OK);
    CHECK(load({7, 1, INT_MAX - 1}) == FGSS_OK);
    CHECK(load({7, 1, INT_MAX}) == FGSS_ERR_UNSUPPORTED);
    CHECK(load({7, 1, INT_MAX, 8, 2, 3, 5}) == FGSS_ERR_UNSUPPORTED);
    load({7, 2, -5, -1});
    load({7, 3, 9, 4, 1, 8, 2, 2, 2});
    load({INT_MIN, 2, 0, 1, INT_MAX, 1, 3});
    // seeded mutations of the fixture: overwritten words (limits, small and random
    // values), truncations, spliced copies of a range
    const int32_t specials[] = {0, 1, -1, 2, 88, 255, 65535, 65536, INT_MAX, INT_MIN, INT_MAX - 1};
    for (int m = 0; m < 4000; m++) {
        std::vector<int32_t> v = fix;
        const int kind = (int)(next_rand() % 4);
        if (kind == 0 || kind == 1) {
            const int edits = 1 + (int)(next_rand() % 3);
            for (int e = 0; e < edits && !v.empty(); e++) {
                const size_t at = next_rand() % v.size();
                v[at] = kind == 0 ? specials[next_rand() % (sizeof(specials) / sizeof(specials[0]))]
                                  : (int32_t)(uint32_t)next_rand();
            }
        } else if (kind == 2) {
            v.resize(next_rand() % (v.size() + 1));
        } else {
            const size_t a = next_rand() % v.size(), b = a + next_rand() % (v.size() - a);
            const size_t at = next_rand() % (v.size() + 1);
            std::vector<int32_t> piece(v.begin() + (long)a, v.begin() + (long)b);
            v.insert(v.begin() + (long)at, piece.begin(), piece.end());
        }
        load(v);
    }
}

static void check_text(const std::vector<char> &fixtext, const std::string &dir)
{
    std::vector<std::string> texts = {
        std::string(fixtext.begin(), fixtext.end()),
        "",
        "\n\n\n",
        "ab\ncd\n",
        "no trailing newline",
        "\nLeading and Trailing\n\n",
        "CRLF line\r\nanother one\r\n",
        "MiXeD CaSe ABC abc AbC\n",
        std::string(40, 'a') + "\n",
        std::string(5000, 'x') + "\n" + std::string(6000, 'y') + "\n",
        std::string("nul\0inside\0\0 the text\n\0\0\0\n", 27),
        "\xff\xfe high bytes \x80\x81\n",
        "same line\nsame line\nsame line\n",
    };
    for (int kind = 0; kind < 2; kind++) {
        char *g = nullptr;
        int64_t nb = 0;
        CHECK(fgss_generate_text(kind, kind ? 60 : 600, 17, &g, &nb) == FGSS_OK);
        texts.emplace_back(g, (size_t)nb);
        fgss_free(g);
    }
    const std::string probe = "probe line\nsame line\nmixed case abc\nzzzz qqqq\n";
    int n = 0;
    for (const std::string &t : texts) {
        for (int q = 1; q <= 12; q++) {
            fgss_collection *c = nullptr;
            const int rc = fgss_collection_from_text_buffer(t.data(), (int64_t)t.size(), q, &c);
            CHECK(rc == FGSS_OK);
            if (rc) continue;
            walk(c);
            fgss_collection *p = nullptr;
            fgss_collection_info ci;
            fgss_collection_get_info(c, &ci);
            if (fgss_collection_from_text_dict(probe.data(), (int64_t)probe.size(), c, &p) == FGSS_OK) {
                walk(p);
                fgss_collection_free(p);
            }
            // write, reload, compare
            const std::string bin = dir + "/c" + std::to_string(n) + ".bin";
            CHECK(fgss_collection_write_bin(c, bin.c_str(), (bin + ".loc").c_str()) == FGSS_OK);
            fgss_collection *r = nullptr;
            CHECK(fgss_collection_from_bin(bin.c_str(), &r) == FGSS_OK);
            if (r) {
                const int32_t *a0, *a1, *b0, *b1;
                const int64_t *ap, *bp;
                fgss_collection_arrays(c, &a0, &a1, &ap, nullptr);
                fgss_collection_arrays(r, &b0, &b1, &bp, nullptr);
                fgss_collection_info ri;
                fgss_collection_get_info(r, &ri);
                CHECK(ri.n_sets == ci.n_sets && ri.n_tokens == ci.n_tokens);
                for (int64_t i = 0; i < ci.n_sets && i < ri.n_sets; i++) CHECK(a0[i] == b0[i] && a1[i] == b1[i]);
                fgss_collection_free(r);
            }
            fgss_collection_free(c);
            n++;
        }
    }
    fgss_collection *c = nullptr;
    CHECK(fgss_collection_from_text_buffer("abc\n", 4, 0, &c) == FGSS_ERR_ARG);
    CHECK(fgss_collection_from_text_buffer("abc\n", 4, 1025, &c) == FGSS_ERR_ARG);
    CHECK(fgss_collection_from_bin((dir + "/does-not-exist.bin").c_str(), &c) == FGSS_ERR_IO);
    CHECK(fgss_collection_from_text((dir + "/does-not-exist.txt").c_str(), 3, &c) == FGSS_ERR_IO);
}

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: %s <fixture .bin> <fixture text> <scratch dir>\n", argv[0]);
        return 2;
    }
    const std::vector<char> raw = read_bytes(argv[1]), text = read_bytes(argv[2]);
    if (raw.empty() || raw.size() % 4 || text.empty()) {
        fprintf(stderr, "cannot read the fixtures\n");
        return 2;
    }
    std::vector<int32_t> fix(raw.size() / 4);
    memcpy(fix.data(), raw.data(), raw.size());
    check_bin(fix);
    check_text(text, argv[3]);
    printf("collection_check: %ld .bin loads, %d failed checks\n", g_loads, g_fail);
    return g_fail ? 1 : 0;
}

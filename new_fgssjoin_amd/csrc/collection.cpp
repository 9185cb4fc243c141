// collection.cpp -- host side of the collection: .bin reader/writer, the q-gram
// text front end, and the synthetic workload generator.
//
// Reference counterparts:
//   .bin reader   ppjoin_format, src/fgssjoin/data.cu:31-126 (record order reversed, :99-112)
//   tokenizer     src/io.cpp:8-26 (non-empty lines), src/preprocess/string/data.cpp:36-40
//                 (A-Z -> a-z), src/preprocess/string/qgram.cpp:20-49 (q-grams suffixed with
//                 their occurrence number), qgram.cpp:9-13 + src/util.cpp:146-157 (64-bit sdbm)
//   reordering    src/preprocess_cli.cpp:266-340 states the intent (rank by document
//                 frequency); the exact rule -- rank by (frequency, first occurrence),
//                 tokens ascending, records by size ascending, ids = 1-based line numbers --
//                 is the one that reproduces test/data/string/small.txt.3gram.bin{,.loc}.
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string_view>
#include <thread>
#include <unordered_map>

#include "fgss_internal.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512];

int fgss_fail(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return status;
}

extern "C" const char *fgss_last_error(void) { return g_err; }
extern "C" const char *fgss_version(void) { return "fgssjoin-mi355x 0.1"; }
extern "C" void fgss_free(void *p) { free(p); }

static unsigned worker_count()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    if (n > 16) n = 16;
    return n;
}

template <class F>
static void parallel_for(int64_t n, int64_t chunk, F &&f)
{
    unsigned nt = worker_count();
    if (n <= chunk || nt == 1) { for (int64_t i = 0; i < n; i++) f(i); return; }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; k++)
        th.emplace_back([&] {
            for (;;) {
                int64_t a = next.fetch_add(chunk);
                if (a >= n) break;
                int64_t b = std::min(n, a + chunk);
                for (int64_t i = a; i < b; i++) f(i);
            }
        });
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------
// collection helpers
// ---------------------------------------------------------------------------
int fgss_finish_collection(fgss_collection *c)
{
    const int64_t n = (int64_t)c->len.size();
    c->pos.resize(n);
    int64_t p = 0;
    int32_t mx = 0;
    bool desc = true;
    for (int64_t i = 0; i < n; i++) {
        c->pos[i] = p;
        p += c->len[i];
        mx = std::max(mx, c->len[i]);
        if (i > 0 && c->len[i] > c->len[i - 1]) desc = false;
    }
    // num_terms = max token + 1 (data.cu:77-78), in 64 bits: a token id of 2^31-1
    // would make the int32 count wrap
    int64_t terms = 0;
    for (int32_t v : c->tok) terms = std::max<int64_t>(terms, (int64_t)v + 1);
    c->max_len = mx;
    c->sorted_desc = desc;
    if (terms > (int64_t)INT32_MAX) {
        c->n_terms = INT32_MAX;
        return fgss_fail(FGSS_ERR_UNSUPPORTED, "token id %lld: ids must be below 2^31-1", (long long)(terms - 1));
    }
    c->n_terms = (int32_t)terms;
    return FGSS_OK;
}

extern "C" void fgss_collection_free(fgss_collection *c)
{
    if (!c) return;
    fgss_device_release(c);
    delete c;
}

extern "C" int fgss_collection_get_info(const fgss_collection *c, fgss_collection_info *info)
{
    if (!c || !info) return fgss_fail(FGSS_ERR_ARG, "null argument");
    info->n_sets = (int64_t)c->len.size();
    info->n_tokens = (int64_t)c->tok.size();
    info->n_terms = c->n_terms;
    info->max_len = c->max_len;
    info->average_size = info->n_sets ? (double)info->n_tokens / (double)info->n_sets : 0.0;
    info->sorted_desc = c->sorted_desc ? 1 : 0;
    info->on_device = c->dev != nullptr;
    return FGSS_OK;
}

extern "C" int fgss_collection_arrays(const fgss_collection *c, const int32_t **id, const int32_t **len,
                                      const int64_t **pos, const int32_t **tokens)
{
    if (!c) return fgss_fail(FGSS_ERR_ARG, "null collection");
    if (id) *id = c->id.data();
    if (len) *len = c->len.data();
    if (pos) *pos = c->pos.data();
    if (tokens) *tokens = c->tok.data();
    return FGSS_OK;
}

// ---------------------------------------------------------------------------
// .bin reader: ppjoin_format (src/fgssjoin/data.cu:31-126)
// ---------------------------------------------------------------------------
extern "C" int fgss_collection_from_bin_buffer(const int32_t *ints, int64_t n_ints, fgss_collection **out)
{
    if (!out || (!ints && n_ints > 0) || n_ints < 0) return fgss_fail(FGSS_ERR_ARG, "null argument");
    std::vector<int64_t> starts;
    int64_t i = 0, ntok = 0;
    while (i < n_ints) {
        if (i + 2 > n_ints) return fgss_fail(FGSS_ERR_FORMAT, "Reading error.");          // data.cu:63-69
        int32_t sz = ints[i + 1];
        if (sz < 0 || i + 2 + (int64_t)sz > n_ints) return fgss_fail(FGSS_ERR_FORMAT, "Reading error.");  // :72-85
        starts.push_back(i);
        ntok += sz;
        i += 2 + (int64_t)sz;
    }
    auto *c = new fgss_collection();
    const int64_t n = (int64_t)starts.size();
    c->id.resize(n);
    c->len.resize(n);
    c->tok.resize(ntok);
    int64_t p = 0;
    for (int64_t k = 0; k < n; k++) {           // reversed file order, data.cu:99-112
        int64_t s = starts[n - 1 - k];
        c->id[k] = ints[s];
        c->len[k] = ints[s + 1];
        // (an empty record copies nothing: tok may still be empty, its data() null)
        if (ints[s + 1] > 0) memcpy(c->tok.data() + p, ints + s + 2, (size_t)ints[s + 1] * sizeof(int32_t));
        p += ints[s + 1];
    }
    const int rc = fgss_finish_collection(c);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return FGSS_OK;
}

static int read_file(const char *path, std::vector<char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return fgss_fail(FGSS_ERR_IO, "File \"%s\" does not exist.", path);   // fgssjoin_cli.cu:26-29
    if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return fgss_fail(FGSS_ERR_IO, "File error"); }
    long sz = ftell(f);
    if (sz < 0) { fclose(f); return fgss_fail(FGSS_ERR_IO, "File error"); }
    rewind(f);
    buf.resize((size_t)sz);
    if (sz > 0 && fread(buf.data(), 1, (size_t)sz, f) != (size_t)sz) {
        fclose(f);
        return fgss_fail(FGSS_ERR_IO, "File error");
    }
    fclose(f);
    return FGSS_OK;
}

extern "C" int fgss_collection_from_bin(const char *path, fgss_collection **out)
{
    if (!path || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    std::vector<char> buf;
    int rc = read_file(path, buf);
    if (rc) return rc;
    // Trailing bytes that do not form a whole int32 are dropped: after the last
    // record they end the reference's read loop (data.cu:56-60), and inside a
    // record the buffer parser reports the short read as "Reading error.".
    int64_t n_ints = (int64_t)(buf.size() / 4);
    std::vector<int32_t> ints((size_t)n_ints);
    if (n_ints) memcpy(ints.data(), buf.data(), (size_t)n_ints * 4);
    return fgss_collection_from_bin_buffer(ints.data(), n_ints, out);
}

extern "C" int fgss_collection_write_bin(const fgss_collection *c, const char *path, const char *loc_path)
{
    if (!c || !path) return fgss_fail(FGSS_ERR_ARG, "null argument");
    FILE *f = fopen(path, "wb");
    if (!f) return fgss_fail(FGSS_ERR_IO, "cannot write \"%s\"", path);
    FILE *fl = nullptr;
    if (loc_path) {
        if (c->loc.size() != c->tok.size()) { fclose(f); return fgss_fail(FGSS_ERR_ARG, "collection has no q-gram offsets"); }
        fl = fopen(loc_path, "wb");
        if (!fl) { fclose(f); return fgss_fail(FGSS_ERR_IO, "cannot write \"%s\"", loc_path); }
    }
    const int64_t n = (int64_t)c->len.size();
    bool ok = true;
    for (int64_t k = n - 1; k >= 0 && ok; k--) {   // file order = internal order reversed
        int32_t hdr[2] = {c->id[k], c->len[k]};
        ok = fwrite(hdr, 4, 2, f) == 2;
        if (ok && c->len[k]) ok = fwrite(c->tok.data() + c->pos[k], 4, (size_t)c->len[k], f) == (size_t)c->len[k];
        if (ok && fl && c->len[k]) ok = fwrite(c->loc.data() + c->pos[k], 4, (size_t)c->len[k], fl) == (size_t)c->len[k];
    }
    if (fclose(f) != 0) ok = false;
    if (fl && fclose(fl) != 0) ok = false;
    return ok ? FGSS_OK : fgss_fail(FGSS_ERR_IO, "write failed");
}

// ---------------------------------------------------------------------------
// q-gram text front end
// ---------------------------------------------------------------------------

// sdbm over gram bytes then the decimal occurrence number, stopping at a NUL
// byte as the reference's C-string walk does (src/util.cpp:146-157).
static inline uint64_t sdbm_step(uint64_t h, unsigned c) { return (uint64_t)c + (h << 6) + (h << 16) - h; }

static uint64_t gram_hash(const unsigned char *g, int q, int occ)
{
    uint64_t h = 0;
    for (int i = 0; i < q; i++) {
        if (g[i] == 0) return h;
        h = sdbm_step(h, g[i]);
    }
    char num[16];
    int nd = snprintf(num, sizeof(num), "%d", occ);
    for (int i = 0; i < nd; i++) h = sdbm_step(h, (unsigned char)num[i]);
    return h;
}

struct HashRef { uint64_t h; uint32_t idx; uint32_t pad; };

// Tokenize text.  Without dict, token ids are ranks by (document frequency, first
// occurrence) and the collection keeps its dictionary (hash -> id).  With dict, the
// q-grams take dict's ids; q-grams dict has never seen get new ids from dict's
// n_terms on, in order of first occurrence (they can match no record of dict).
static int tokenize_text(const char *text, int64_t nbytes, int q, const fgss_collection *dict,
                         fgss_collection **out)
{
    if (!out || (!text && nbytes > 0) || nbytes < 0) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (q < 1 || q > 1024) return fgss_fail(FGSS_ERR_ARG, "q must be in 1..1024");   // preprocess_cli.cpp:86
    // non-empty lines (io.cpp:16-19)
    std::vector<int64_t> ls, ll;
    int64_t s0 = 0;
    for (int64_t i = 0; i <= nbytes; i++) {
        if (i == nbytes || text[i] == '\n') {
            if (i > s0) { ls.push_back(s0); ll.push_back(i - s0); }
            s0 = i + 1;
        }
    }
    const int64_t nl = (int64_t)ls.size();
    std::vector<int64_t> roff(nl + 1, 0);
    for (int64_t r = 0; r < nl; r++) roff[r + 1] = roff[r] + std::max<int64_t>(0, ll[r] - q + 1);
    const int64_t total = roff[nl];
    if (total >= (int64_t)UINT32_MAX) return fgss_fail(FGSS_ERR_UNSUPPORTED, "more than 2^32-1 q-grams");
    std::vector<HashRef> hr((size_t)total);
    std::vector<int32_t> gpos((size_t)total);

    // per-record q-grams with occurrence suffix (qgram.cpp:20-49)
    parallel_for(nl, 256, [&](int64_t r) {
        const int64_t L = ll[r], ng = roff[r + 1] - roff[r];
        if (ng <= 0) return;
        std::string low((size_t)L, '\0');
        for (int64_t i = 0; i < L; i++) {
            unsigned char ch = (unsigned char)text[ls[r] + i];
            low[(size_t)i] = (char)((ch >= 'A' && ch <= 'Z') ? ch - ('Z' - 'z') : ch);   // data.cpp:36-40
        }
        const unsigned char *u = (const unsigned char *)low.data();
        std::vector<int> occ((size_t)ng, 0);
        if (q <= 7) {
            std::vector<std::pair<uint64_t, uint32_t>> keys((size_t)ng);
            for (int64_t i = 0; i < ng; i++) {
                uint64_t k = 0;
                for (int j = 0; j < q; j++) k = (k << 8) | u[i + j];
                keys[(size_t)i] = {k, (uint32_t)i};
            }
            std::sort(keys.begin(), keys.end());
            for (size_t i = 1; i < keys.size(); i++)
                if (keys[i].first == keys[i - 1].first) occ[keys[i].second] = occ[keys[i - 1].second] + 1;
        } else {
            std::unordered_map<std::string_view, int> seen;
            for (int64_t i = 0; i < ng; i++) {
                std::string_view g(low.data() + i, (size_t)q);
                auto it = seen.find(g);
                if (it == seen.end()) seen.emplace(g, 0);
                else occ[(size_t)i] = ++it->second;
            }
        }
        for (int64_t i = 0; i < ng; i++) {
            int64_t f = roff[r] + i;
            hr[(size_t)f] = {gram_hash(u + i, q, occ[(size_t)i]), (uint32_t)f, 0};
            gpos[(size_t)f] = (int32_t)i;
        }
    });

    // sort by (hash, flat index): bucket by the top byte, sort buckets in parallel
    {
        std::vector<int64_t> cnt(257, 0);
        for (const auto &x : hr) cnt[(x.h >> 56) + 1]++;
        for (int k = 0; k < 256; k++) cnt[k + 1] += cnt[k];
        std::vector<HashRef> tmp((size_t)total);
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (const auto &x : hr) tmp[(size_t)fill[x.h >> 56]++] = x;
        parallel_for(256, 1, [&](int64_t b) {
            std::sort(tmp.begin() + cnt[b], tmp.begin() + cnt[b + 1],
                      [](const HashRef &a, const HashRef &c) { return a.h != c.h ? a.h < c.h : a.idx < c.idx; });
        });
        hr.swap(tmp);
    }
    // distinct hashes: (df, first occurrence) -> rank = token id
    std::vector<int64_t> gstart;
    for (int64_t i = 0; i < total; i++)
        if (i == 0 || hr[(size_t)i].h != hr[(size_t)i - 1].h) gstart.push_back(i);
    const int64_t nd = (int64_t)gstart.size();
    gstart.push_back(total);
    std::vector<int64_t> gorder((size_t)nd);
    for (int64_t g = 0; g < nd; g++) gorder[(size_t)g] = g;
    std::sort(gorder.begin(), gorder.end(), [&](int64_t a, int64_t b) {
        int64_t da = gstart[a + 1] - gstart[a], db = gstart[b + 1] - gstart[b];
        if (da != db) return da < db;
        return hr[(size_t)gstart[a]].idx < hr[(size_t)gstart[b]].idx;
    });
    std::vector<int32_t> tid((size_t)total);
    std::vector<uint64_t> dict_hash;
    std::vector<int32_t> dict_id;
    int64_t n_terms = nd;
    if (!dict) {
        parallel_for(nd, 4096, [&](int64_t k) {
            int64_t g = gorder[(size_t)k];
            for (int64_t i = gstart[g]; i < gstart[g + 1]; i++) tid[hr[(size_t)i].idx] = (int32_t)k;
        });
        dict_hash.resize((size_t)nd);
        dict_id.resize((size_t)nd);
        for (int64_t k = 0; k < nd; k++) {   // groups are in hash order
            dict_hash[(size_t)k] = hr[(size_t)gstart[k]].h;
            dict_id[(size_t)k] = 0;
        }
        for (int64_t k = 0; k < nd; k++) dict_id[(size_t)gorder[(size_t)k]] = (int32_t)k;
    } else {
        // unseen q-grams by first occurrence
        std::vector<std::pair<uint32_t, int64_t>> unseen;
        std::vector<int32_t> gid((size_t)nd, -1);
        for (int64_t g = 0; g < nd; g++) {
            const uint64_t h = hr[(size_t)gstart[g]].h;
            auto it = std::lower_bound(dict->dict_hash.begin(), dict->dict_hash.end(), h);
            if (it != dict->dict_hash.end() && *it == h) gid[(size_t)g] = dict->dict_id[(size_t)(it - dict->dict_hash.begin())];
            else unseen.push_back({hr[(size_t)gstart[g]].idx, g});
        }
        std::sort(unseen.begin(), unseen.end());
        for (size_t k = 0; k < unseen.size(); k++) gid[(size_t)unseen[k].second] = dict->n_terms + (int32_t)k;
        n_terms = (int64_t)dict->n_terms + (int64_t)unseen.size();
        if (n_terms >= INT32_MAX) return fgss_fail(FGSS_ERR_UNSUPPORTED, "too many distinct q-grams");
        parallel_for(nd, 4096, [&](int64_t g) {
            for (int64_t i = gstart[g]; i < gstart[g + 1]; i++) tid[hr[(size_t)i].idx] = gid[(size_t)g];
        });
    }
    std::vector<HashRef>().swap(hr);

    // records by size ascending (stable in line order); internal order = reversed
    std::vector<int64_t> recs;
    for (int64_t r = 0; r < nl; r++) if (roff[r + 1] > roff[r]) recs.push_back(r);
    std::stable_sort(recs.begin(), recs.end(),
                     [&](int64_t a, int64_t b) { return roff[a + 1] - roff[a] < roff[b + 1] - roff[b]; });
    const int64_t n = (int64_t)recs.size();
    auto *c = new fgss_collection();
    c->id.resize(n);
    c->len.resize(n);
    for (int64_t k = 0; k < n; k++) {
        int64_t r = recs[(size_t)(n - 1 - k)];
        c->id[k] = (int32_t)(r + 1);
        c->len[k] = (int32_t)(roff[r + 1] - roff[r]);
    }
    (void)fgss_finish_collection(c);   // offsets and lengths (no tokens yet; n_terms is set below)
    c->tok.resize((size_t)total);
    c->loc.resize((size_t)total);
    parallel_for(n, 256, [&](int64_t k) {
        int64_t r = recs[(size_t)(n - 1 - k)], ng = roff[r + 1] - roff[r];
        std::vector<std::pair<int32_t, int32_t>> v((size_t)ng);
        for (int64_t i = 0; i < ng; i++) v[(size_t)i] = {tid[(size_t)(roff[r] + i)], gpos[(size_t)(roff[r] + i)]};
        std::sort(v.begin(), v.end());
        for (int64_t i = 0; i < ng; i++) {
            c->tok[(size_t)(c->pos[k] + i)] = v[(size_t)i].first;
            c->loc[(size_t)(c->pos[k] + i)] = v[(size_t)i].second;
        }
    });
    c->n_terms = (int32_t)n_terms;
    c->q = q;
    c->dict_hash.swap(dict_hash);
    c->dict_id.swap(dict_id);
    *out = c;
    return FGSS_OK;
}

extern "C" int fgss_collection_from_text_buffer(const char *text, int64_t nbytes, int q, fgss_collection **out)
{
    return tokenize_text(text, nbytes, q, nullptr, out);
}

extern "C" int fgss_collection_from_text_dict(const char *text, int64_t nbytes, const fgss_collection *dict,
                                              fgss_collection **out)
{
    if (!dict || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (dict->q < 1 || dict->dict_hash.empty())
        return fgss_fail(FGSS_ERR_ARG, "the dictionary collection was not built from text");
    return tokenize_text(text, nbytes, dict->q, dict, out);
}

extern "C" int fgss_collection_from_text(const char *path, int q, fgss_collection **out)
{
    if (!path || !out) return fgss_fail(FGSS_ERR_ARG, "null argument");
    std::vector<char> buf;
    int rc = read_file(path, buf);
    if (rc) return rc;
    return fgss_collection_from_text_buffer(buf.data(), (int64_t)buf.size(), q, out);
}

// ---------------------------------------------------------------------------
// synthetic workload generator (seeded, portable: splitmix64 + integer CDFs)
// ---------------------------------------------------------------------------
namespace {
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }                   // modulo bias is irrelevant here
    int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)below((uint64_t)(hi - lo + 1)); }
};

// English letter frequencies (per 10,000), a..z
const int kLetterW[26] = {817, 149, 278, 425, 1270, 223, 202, 609, 697, 15, 77, 403, 241,
                          675, 751, 193, 10, 599, 633, 906, 276, 98, 236, 15, 197, 7};

struct Vocab {
    std::vector<std::string> words;
    std::vector<uint64_t> cdf;      // Zipf(s=1) cumulative weights, scaled to 2^62
    int lcdf[27];
    char letter(Rng &r) const
    {
        int u = (int)r.below((uint64_t)lcdf[26]);
        int k = 0;
        while (lcdf[k + 1] <= u) k++;
        return (char)('a' + k);
    }
    const std::string &zipf(Rng &r) const
    {
        uint64_t u = r.next() >> 2;
        size_t k = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (k >= words.size()) k = words.size() - 1;
        return words[k];
    }
};

Vocab make_vocab(uint64_t seed, int n_words)
{
    Vocab v;
    v.lcdf[0] = 0;
    for (int k = 0; k < 26; k++) v.lcdf[k + 1] = v.lcdf[k] + kLetterW[k];
    Rng r(seed ^ 0x7f4a7c159e3779b9ull);
    v.words.reserve((size_t)n_words);
    for (int i = 0; i < n_words; i++) {
        int L = 3;                                      // 3-12 chars, geometric (p = 0.8 per extra char)
        while (L < 12 && r.below(10) < 8) L++;
        std::string w;
        for (int k = 0; k < L; k++) w.push_back(v.letter(r));
        v.words.push_back(w);
    }
    // Zipf(1): weight 1/rank; integer CDF so both machines draw identical words
    std::vector<double> acc((size_t)n_words);
    double h = 0.0;
    for (int i = 0; i < n_words; i++) { h += 1.0 / (double)(i + 1); acc[(size_t)i] = h; }
    v.cdf.resize((size_t)n_words);
    const double scale = (double)(1ull << 62);
    for (int i = 0; i < n_words; i++) v.cdf[(size_t)i] = (uint64_t)(acc[(size_t)i] / h * scale);
    v.cdf[(size_t)n_words - 1] = 1ull << 62;
    return v;
}

void edit(std::string &s, int n_edits, const Vocab &v, Rng &r)
{
    for (int e = 0; e < n_edits; e++) {
        int op = (int)r.below(3);
        if (s.empty()) op = 1;
        // draw the position before the letter: one fixed order on every compiler
        if (op == 0) {                                                   // substitute
            size_t at = (size_t)r.below(s.size());
            s[at] = v.letter(r);
        } else if (op == 1) {                                            // insert
            size_t at = (size_t)r.below(s.size() + 1);
            char ch = v.letter(r);
            s.insert(s.begin() + (long)at, ch);
        } else {                                                         // delete
            s.erase(s.begin() + (long)r.below(s.size()));
        }
    }
}
}  // namespace

extern "C" int fgss_generate_text(int kind, int64_t n_records, uint64_t seed, char **text, int64_t *nbytes)
{
    if (!text || !nbytes || n_records < 0) return fgss_fail(FGSS_ERR_ARG, "null argument");
    if (kind != FGSS_GEN_TITLES && kind != FGSS_GEN_LONG) return fgss_fail(FGSS_ERR_ARG, "unknown generator kind");
    const Vocab v = make_vocab(seed, 50000);
    Rng r(seed);
    const bool titles = kind == FGSS_GEN_TITLES;
    const double dup = titles ? 0.10 : 0.05;
    const int emin = 1, emax = titles ? 3 : 10;
    const uint64_t dup_cut = (uint64_t)(dup * 1e9);
    std::vector<std::string> recs;
    recs.reserve((size_t)n_records);
    size_t bytes = 0;
    for (int64_t i = 0; i < n_records; i++) {
        std::string s;
        if (i > 0 && r.below(1000000000ull) < dup_cut) {
            s = recs[(size_t)r.below((uint64_t)i)];
            edit(s, (int)r.range(emin, emax), v, r);
        } else if (titles) {
            int nw = (int)r.range(4, 15);
            for (int w = 0; w < nw; w++) { if (w) s.push_back(' '); s += v.zipf(r); }
            while (s.size() < 20) { s.push_back(' '); s += v.zipf(r); }
            if (s.size() > 120) {
                size_t cut = s.rfind(' ', 120);
                s.resize(cut != std::string::npos && cut >= 20 ? cut : 120);
            }
        } else {
            size_t target = (size_t)r.range(200, 1200);
            while (s.size() < target) { if (!s.empty()) s.push_back(' '); s += v.zipf(r); }
            size_t cut = s.rfind(' ', target);
            s.resize(cut != std::string::npos && cut >= 200 ? cut : target);
        }
        bytes += s.size() + 1;
        recs.push_back(std::move(s));
    }
    char *buf = (char *)malloc(bytes ? bytes : 1);
    if (!buf) return fgss_fail(FGSS_ERR_NOMEM, "out of host memory");
    size_t o = 0;
    for (const auto &s : recs) { memcpy(buf + o, s.data(), s.size()); o += s.size(); buf[o++] = '\n'; }
    *text = buf;
    *nbytes = (int64_t)bytes;
    return FGSS_OK;
}

// cli.cpp -- the fgssjoin command-line tool over the C ABI.
//
// Forms:
//   fgssjoin <ppjoin bin file> <threshold> [verbose|debug]      (src/fgssjoin_cli.cu:20-49)
//   fgssjoin -f <text file> -q <q> -t <threshold> [verbose|debug] [-o <file.bin>]
//                                                               (README.md:12-15, 45-51)
// stdout: one "id id similarity" line per result pair (src/fgssjoin/block.cu:193);
// stderr: the reference's counters and timings (fgssjoin_cli.cu:55-108, block.cu:209-212).
// With -o the tokenized collection is also written as <file.bin> and <file.bin>.loc.
// Text is tokenized on the GPU for q <= 8 (identical result), on the host otherwise or
// with --host-tokenizer.  debug also dumps the inverted index (fgssjoin_cli.cu:93-94).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "fgssjoin.h"

static double now_ms()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec * 1e3 + (double)tv.tv_usec / 1e3;
}

// whole file into buf; FGSS_ERR_IO when it cannot be read
static int read_all(const char *path, std::vector<char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "File error\n");
        return FGSS_ERR_IO;
    }
    char chunk[1 << 16];
    size_t k;
    while ((k = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + k);
    fclose(f);
    return FGSS_OK;
}

// the handles of a run, released on every return path
struct Handles {
    fgss_collection *c = nullptr;
    fgss_pairs *p = nullptr;
    ~Handles()
    {
        fgss_pairs_free(p);
        fgss_collection_free(c);
    }
};

static int usage(const char *argv0)
{
    fprintf(stdout, "Usage:\n./%s <ppjoin bin file> <threshold> [verbose]\n", argv0);
    fprintf(stdout, "       %s -f <text file> -q <q> -t <threshold> [verbose|debug] [-o <out.bin>] [--host-tokenizer]\n",
            argv0);
    return 1;
}

int main(int argc, char **argv)
{
    const char *path = nullptr, *tstr = nullptr, *outbin = nullptr;
    int q = 0;
    bool text = false, verbose = false, debug = false, host_tok = false;
    if (argc < 3) return usage(argv[0]);
    if (argv[1][0] == '-' && argv[1][1] != '\0') {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "-f" && i + 1 < argc) { path = argv[++i]; text = true; }
            else if (a == "-q" && i + 1 < argc) q = atoi(argv[++i]);
            else if (a == "-t" && i + 1 < argc) tstr = argv[++i];
            else if (a == "-o" && i + 1 < argc) outbin = argv[++i];
            else if (a == "--host-tokenizer") host_tok = true;
            else if (a == "verbose") verbose = true;
            else if (a == "debug") debug = true;
            else return usage(argv[0]);
        }
        if (!path || !tstr || q < 1 || q > 1024) return usage(argv[0]);
    } else {
        path = argv[1];
        tstr = argv[2];
        for (int i = 3; i < argc && i < 5; i++) {   // fgssjoin_cli.cu:37-49
            if (strcmp(argv[i], "verbose") == 0) verbose = true;
            if (strcmp(argv[i], "debug") == 0) debug = true;
        }
    }
    if (access(path, F_OK) == -1) {   // fgssjoin_cli.cu:26-29
        fprintf(stderr, "File \"%s\" does not exist.\n", path);
        return 1;
    }
    float threshold = (float)atof(tstr);   // fgssjoin_cli.cu:31
    if (threshold <= 0.0f || threshold >= 1.0) {
        fprintf(stderr, "Threshold must be a float in the range (0.0..1.0).\n");
        return 1;
    }

    fprintf(stderr, "Document: %s\n", path);
    fprintf(stderr, "Algorithm: %s\n", "fgssjoin");
    fprintf(stderr, "Threshold: Jaccard %g\n", threshold);
    fprintf(stderr, "... LOADING DATASET ...\n");
    Handles h;
    fgss_collection *&c = h.c;
    int rc;
    if (text && q <= 8 && !host_tok) {   // the device tokenizer packs q-grams of up to 8 bytes
        std::vector<char> buf;
        rc = read_all(path, buf);
        if (rc == FGSS_OK) rc = fgss_collection_from_text_device(buf.data(), (int64_t)buf.size(), q, &c);
    } else {
        rc = text ? fgss_collection_from_text(path, q, &c) : fgss_collection_from_bin(path, &c);
    }
    if (rc != FGSS_OK) {
        const char *msg = fgss_last_error();
        if (rc == FGSS_ERR_FORMAT) fprintf(stdout, "%s\n", msg);   // data.cu:67,83 print to stdout
        else fprintf(stderr, "%s\n", msg);
        return 1;
    }
    if (outbin) {
        std::string loc = std::string(outbin) + ".loc";
        if (fgss_collection_write_bin(c, outbin, text ? loc.c_str() : nullptr) != FGSS_OK) {
            fprintf(stderr, "%s\n", fgss_last_error());
            return 1;
        }
    }
    fgss_collection_info info;
    fgss_collection_get_info(c, &info);
    fprintf(stderr, "# Records: %lld\n", (long long)info.n_sets);
    fprintf(stderr, "# Terms: %d\n", info.n_terms);
    fprintf(stderr, "# Tokens: %lld\n", (long long)info.n_tokens);
    fprintf(stderr, "# Average size: %g\n", info.average_size);
    if (debug) {   // fgssjoin_cli.cu:66-74
        const int32_t *id, *len, *tok;
        const int64_t *pos;
        fgss_collection_arrays(c, &id, &len, &pos, &tok);
        for (int64_t i = 0; i < info.n_sets; i++) {
            fprintf(stdout, "id: [%u]\n", (unsigned)id[i]);
            fprintf(stdout, "size: [%u]\n", (unsigned)len[i]);
            for (int32_t j = 0; j < len[i]; j++) fprintf(stdout, "(%u) ", (unsigned)tok[pos[i] + j]);
            fprintf(stdout, "\n\n");
        }
    }

    fprintf(stderr, "=== BEGIN JOIN (TIMER STARTED) ===\n");
    double t0 = now_ms();
    fprintf(stderr, "# Preparing and transfering...\n");
    double t00 = now_ms();
    rc = fgss_collection_upload(c);
    double t01 = now_ms();
    if (rc != FGSS_OK) {
        fprintf(stderr, "%s\n", fgss_last_error());
        return 1;
    }
    fprintf(stderr, "  - Time spent: %g ms\n", t01 - t00);
    if (debug && fgss_collection_print_index(c, threshold, stderr) != FGSS_OK) {   // fgssjoin_cli.cu:93-94
        fprintf(stderr, "%s\n", fgss_last_error());
        return 1;
    }
    fgss_pairs *&p = h.p;
    rc = fgss_selfjoin(c, threshold, nullptr, &p);
    if (rc != FGSS_OK) {
        fprintf(stderr, "%s\n", fgss_last_error());
        return 1;
    }
    fprintf(stderr, "# Inverted Index Entries: %lld\n", (long long)p->stats.index_entries);
    fprintf(stderr, "> Block processing with static index.\n");
    if (verbose) {
        fprintf(stderr, "  - index build: %g ms\n", p->stats.ms_index);
        fprintf(stderr, "  - candidate generation + verification: %g ms\n", p->stats.ms_join);
        fprintf(stderr, "  - postings scanned: %lld\n", (long long)p->stats.postings_scanned);
    }
    double r0 = now_ms();
    fgss_pairs_print(p, stdout);
    double r1 = now_ms();
    fprintf(stderr, "# Candidates: %lld\n", (long long)p->stats.candidates);
    fprintf(stderr, "# Intersections: %lld\n", (long long)p->stats.intersections);
    fprintf(stderr, "# Results: %lld\n", (long long)p->stats.results);
    fprintf(stderr, "  - Time rendering: %g ms\n", r1 - r0);
    double t1 = now_ms();
    fprintf(stderr, "===  END JOIN (TIMER STOPPED)  ===\n");
    fprintf(stderr, "Total Running Time: %g s\n", (t1 - t0) / 1000.0);
    return 0;
}

// Driver around the reference's own host tokenizer, compiled from the sources
// under /root/reference (src/io.cpp, src/util.cpp, src/preprocess/string/{qgram,data}.cpp)
// by oracle/Makefile into oracle/_ref/libref_tok.so.  TEST INFRASTRUCTURE ONLY:
// it checks the oracle's q-gram / occurrence-suffix / sdbm restatement.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "io.hpp"
#include "data.hpp"
#include "qgram.hpp"

// Reference calls: io_getlines (io.cpp:8-26), dat_proc_data (data.cpp:74-90),
// qg_get_records (qgram.cpp:56-71), qg_get_sets (qgram.cpp:78-95).
// Returns the number of hashes; *hashes (n) and *rec_len (one per non-empty line)
// are malloc'd.
extern "C" int64_t ref_qgram_hashes(const char *path, int q, uint64_t **hashes, int64_t **rec_len, int64_t *nrec)
{
    vector<string> lines = io_getlines(string(path));
    dat_proc_data(lines);
    vector<vector<string> > recs = qg_get_records(lines, q);
    vector<vector<unsigned long> > sets = qg_get_sets(recs);
    int64_t total = 0;
    for (size_t i = 0; i < sets.size(); i++) total += (int64_t)sets[i].size();
    *hashes = (uint64_t *)malloc((size_t)(total > 0 ? total : 1) * sizeof(uint64_t));
    *rec_len = (int64_t *)malloc((sets.size() > 0 ? sets.size() : 1) * sizeof(int64_t));
    int64_t k = 0;
    for (size_t i = 0; i < sets.size(); i++) {
        (*rec_len)[i] = (int64_t)sets[i].size();
        for (size_t j = 0; j < sets[i].size(); j++) (*hashes)[k++] = (uint64_t)sets[i][j];
    }
    *nrec = (int64_t)sets.size();
    return total;
}

extern "C" void ref_free(void *p) { free(p); }

// fgss_internal.h -- host-side structures shared by the library's translation units.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fgssjoin.h"

// Set the thread-local error message and return the status.
int fgss_fail(int status, const char *fmt, ...);

struct fgss_device_sets;   // defined in join.hip

// A collection in the reference's internal order (the .bin file order
// reversed, src/fgssjoin/data.cu:99-112).
struct fgss_collection {
    std::vector<int32_t> id;
    std::vector<int32_t> len;
    std::vector<int64_t> pos;       // exclusive scan of len
    std::vector<int32_t> tok;       // concatenated records, each ascending
    std::vector<int32_t> loc;       // q-gram offsets (text front end only), parallel to tok
    std::vector<uint32_t> job;      // batches only: job id per record
    std::vector<uint32_t> jend;     // batches only: per job, one past its last record
    std::vector<uint64_t> dict_hash;  // text front end: q-gram hashes, ascending
    std::vector<int32_t> dict_id;     // ... and their token ids
    int q = 0;                        // text front end: q-gram length
    int32_t n_terms = 0;
    int32_t max_len = 0;
    bool sorted_desc = true;
    fgss_device_sets *dev = nullptr;   // resident device copy + reusable workspace
};

// Fill len/pos/n_terms/max_len/sorted_desc from id/len/tok already in internal order.
// FGSS_ERR_UNSUPPORTED (n_terms clamped) when a token id is 2^31-1 (n_terms would not fit int32).
int fgss_finish_collection(fgss_collection *c);

// Device-side teardown (join.hip).
void fgss_device_release(fgss_collection *c);

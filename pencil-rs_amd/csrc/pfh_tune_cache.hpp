// Batch tunings kept across plan creations (process map) and processes (a
// text file); see pfh_tune_cache.cpp.
#pragma once

#include <cstddef>
#include <string>

#include "../../include/pencil_fft_hip.h"

namespace pfh {

struct TuneRecord {
    int block = 256;
    int max_radix = 8;
    int col_lines = 0;
    int multi = 1;
    int persistent = 1;
    int plane = 1;
    int streams = 1;
    double ms = 0; // the measured forward time (informational)
};

// The key of a batch: the device (architecture name, CU count, LDS bytes per
// workgroup) and every descriptor field the batch planner reads.
std::string batch_tune_key(const char *arch, int cus, size_t lds_limit, const pfh_desc *descs, int count);

// Looks a key up in the process map, then in the cache file; *source is
// PFH_TUNED_PROCESS_CACHE or PFH_TUNED_FILE_CACHE on a hit.
bool tune_cache_lookup(const std::string &key, TuneRecord &out, int *source);
// Records a measured tuning in the process map and appends it to the file.
void tune_cache_store(const std::string &key, const TuneRecord &r);
// Drops one key from the process map (a record that no longer plans).
void tune_cache_forget(const std::string &key);
// Empties the process map (the file is kept).
void tune_cache_clear();

} // namespace pfh

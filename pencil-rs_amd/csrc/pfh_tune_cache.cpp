// Measured batch tunings kept across plan creations and processes.
//
// The batch autotuner times every candidate tuning on the device, which
// costs seconds for a large batch. Its choice depends only on the problems'
// descriptors, the precision and the device (architecture name, CU count,
// LDS per workgroup), so it is kept under a key made of those: in a map for
// the process, and as one text line per measurement in a file
//
//   <key> <block> <max_radix> <col_lines> <multi> <persistent> <plane> <streams> <ms>
//
// that later processes read (the last line of a key wins; lines that do not
// parse are skipped). The file is PFH_TUNE_CACHE when that is set (an empty
// value, "0" or "off" keeps the tunings in the process only), else
// $XDG_CACHE_HOME/pencil-fft-hip/batch_tuning.txt or
// $HOME/.cache/pencil-fft-hip/batch_tuning.txt. The cache is best effort: a
// file that cannot be read or written only means measuring again.
#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "pfh_tune_cache.hpp"

namespace pfh {

namespace {

std::mutex g_mu;
std::map<std::string, TuneRecord> g_mem;

// Bumped when the meaning of a record's fields or the candidate set changes,
// so that records written by another build are not reused.
constexpr const char *kKeyVersion = "pfhbt3";

std::string cache_path()
{
    const char *e = std::getenv("PFH_TUNE_CACHE");
    if (e) {
        if (!*e || !std::strcmp(e, "0") || !std::strcmp(e, "off")) return std::string();
        return std::string(e);
    }
    const char *x = std::getenv("XDG_CACHE_HOME");
    if (x && *x) return std::string(x) + "/pencil-fft-hip/batch_tuning.txt";
    const char *h = std::getenv("HOME");
    if (h && *h) return std::string(h) + "/.cache/pencil-fft-hip/batch_tuning.txt";
    return std::string();
}

bool valid(const TuneRecord &r)
{
    const bool block_ok = r.block >= 64 && r.block <= 1024 && (r.block & (r.block - 1)) == 0;
    const bool radix_ok = r.max_radix >= 2 && r.max_radix <= 64;
    const bool col_ok = r.col_lines >= 0 && r.col_lines <= 256;
    auto flag = [](int v) { return v == 0 || v == 1; };
    return block_ok && radix_ok && col_ok && flag(r.multi) && flag(r.persistent) && flag(r.plane) &&
           (r.streams == 1 || r.streams == 2);
}

// Parses one line of the file; false for anything that is not a record.
bool parse_line(const std::string &line, std::string &key, TuneRecord &r)
{
    const size_t sp = line.find(' ');
    if (sp == std::string::npos || sp == 0) return false;
    key = line.substr(0, sp);
    if (key.compare(0, std::strlen(kKeyVersion), kKeyVersion) != 0) return false;
    double ms = 0;
    int used = 0;
    const int n = std::sscanf(line.c_str() + sp, " %d %d %d %d %d %d %d %lf%n", &r.block, &r.max_radix,
                              &r.col_lines, &r.multi, &r.persistent, &r.plane, &r.streams, &ms, &used);
    if (n != 8) return false;
    for (const char *p = line.c_str() + sp + used; *p; ++p)
        if (*p != ' ' && *p != '\r') return false;
    r.ms = ms;
    return valid(r);
}

bool file_lookup(const std::string &path, const std::string &key, TuneRecord &out)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    bool found = false;
    std::string line, k;
    char buf[4096];
    while (std::fgets(buf, sizeof buf, f)) {
        line += buf;
        if (line.empty() || line.back() != '\n') {
            if (!std::feof(f)) continue; // a long line continues in the next chunk
        } else {
            line.pop_back();
        }
        TuneRecord r;
        if (parse_line(line, k, r) && k == key) {
            out = r;
            found = true;
        }
        line.clear();
    }
    std::fclose(f);
    return found;
}

void make_parent_dirs(const std::string &path)
{
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') (void)::mkdir(path.substr(0, i).c_str(), 0755);
}

void file_append(const std::string &path, const std::string &key, const TuneRecord &r)
{
    make_parent_dirs(path);
    const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC, 0644);
    if (fd < 0) return;
    char tail[160];
    std::snprintf(tail, sizeof tail, " %d %d %d %d %d %d %d %.6f\n", r.block, r.max_radix, r.col_lines, r.multi,
                  r.persistent, r.plane, r.streams, r.ms);
    const std::string line = key + tail;
    // one write per record under an exclusive lock: records of processes
    // that tune at the same time do not interleave
    if (::flock(fd, LOCK_EX) == 0) {
        size_t off = 0;
        while (off < line.size()) {
            const ssize_t w = ::write(fd, line.data() + off, line.size() - off);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) break;
            off += (size_t)w;
        }
        (void)::flock(fd, LOCK_UN);
    }
    ::close(fd);
}

} // namespace

std::string batch_tune_key(const char *arch, int cus, size_t lds_limit, const pfh_desc *descs, int count)
{
    std::string k = kKeyVersion;
    k += ";arch=";
    for (const char *p = arch ? arch : ""; *p; ++p) k += (*p == ' ' || *p == '\t' || *p == '\n') ? '_' : *p;
    char b[128];
    std::snprintf(b, sizeof b, ";cu=%d;lds=%zu;n=%d", cus, lds_limit, count);
    k += b;
    for (int i = 0; i < count; ++i) {
        const pfh_desc &d = descs[i];
        std::snprintf(b, sizeof b, ";k%dp%dd%d:", d.kind, d.precision, d.ndim);
        k += b;
        for (int a = 0; a < d.ndim && a < PFH_MAX_DIMS; ++a) {
            std::snprintf(b, sizeof b, "%s%zu", a ? "x" : "", d.shape[a]);
            k += b;
        }
        k += "e";
        for (int a = 0; a < d.extra_ndim && a < PFH_MAX_DIMS; ++a) {
            std::snprintf(b, sizeof b, "%s%zu", a ? "x" : "", d.extra_shape[a]);
            k += b;
        }
        std::snprintf(b, sizeof b, "s%ut%dm%d", d.selected, d.topology_dims, d.permute_dims);
        k += b;
    }
    return k;
}

bool tune_cache_lookup(const std::string &key, TuneRecord &out, int *source)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mem.find(key);
    if (it != g_mem.end()) {
        out = it->second;
        if (source) *source = PFH_TUNED_PROCESS_CACHE;
        return true;
    }
    const std::string path = cache_path();
    if (path.empty()) return false;
    TuneRecord r;
    if (!file_lookup(path, key, r)) return false;
    g_mem[key] = r;
    out = r;
    if (source) *source = PFH_TUNED_FILE_CACHE;
    return true;
}

void tune_cache_store(const std::string &key, const TuneRecord &r)
{
    if (!valid(r)) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_mem[key] = r;
    const std::string path = cache_path();
    if (!path.empty()) file_append(path, key, r);
}

void tune_cache_forget(const std::string &key)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_mem.erase(key);
}

void tune_cache_clear()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_mem.clear();
}

} // namespace pfh

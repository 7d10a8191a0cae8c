// On-file formats for device pencil arrays at one rank: the single-dataset
// format of the reference's write_mpi / read_mpi (crates/pencil-io/src/
// mpi_io.rs:198-840, header :1492-1930, element codes format.rs:50-126), its
// headerless raw input (raw.rs:24-238) and its metadata catalog
// (catalog.rs:113-300).
//
// File = 96-byte prefix + u64 arrays + payload, all little endian:
//   0 "PENCILIO" | 8 version (1) | 16 header length | 24 commit marker
//   32 N | 40 type code | 48 width | 56 extra rank | 64 payload offset
//   72 payload bytes | 80 writer grid rank | 88 reserved (0)
//   96: global[N], extra[extra rank], writer grid[grid rank], writer perm[N]
//   payload: canonical row-major [extra..., logical spatial...] values.
// The commit marker is INCOMPLETE until the payload is written and flushed.
//
// Device side: the pencil's memory order is permuted to (write) or from
// (read) the canonical order on the GPU with the word-permutation kernel,
// through a device staging buffer; host transfers go through two pinned
// chunks so a chunk's disk write overlaps the next chunk's copy. A read
// stages and validates the whole payload before the destination is written.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pencil_fft_hip.h"
#include "pfh_kernels.hpp"

namespace pfh {
pfh_status set_error(pfh_status st, const char *msg, size_t required, size_t actual, size_t batch, size_t index,
                     int hip_error);
pfh_status clear_error();
} // namespace pfh

using namespace pfh;

namespace {

constexpr size_t kPrefix = 96;
constexpr uint64_t kVersion = 1;
constexpr uint64_t kCommit = 0x434f4d4d49545445ull;     // "ETTIMMOC" little endian
constexpr uint64_t kIncomplete = 0x494e434f4d504c45ull; // "ELPMOCNI"
constexpr uint64_t kMaxHeader = 1024 * 1024;
constexpr uint64_t kMaxRank = 1024;
constexpr size_t kMaxDescriptor = 64 * 1024;
constexpr size_t kChunk = 32u << 20;
const char kMagic[8] = {'P', 'E', 'N', 'C', 'I', 'L', 'I', 'O'};

pfh_status err(pfh_status st, const char *what, size_t actual = 0)
{
    return set_error(st, what, 0, actual, 0, 0, 0);
}

#define IO_HIP(expr)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return set_error(PFH_ERR_DEVICE, hipGetErrorString(e_), 0, 0, 0, 0, (int)e_); \
    } while (0)

// IoElement width per type code (format.rs:50-60, :95-126)
int type_width(uint64_t code)
{
    static const int w[13] = {0, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8, 8, 16};
    return code >= 1 && code <= 12 ? w[code] : 0;
}

void put_u64(uint8_t *b, size_t off, uint64_t v)
{
    for (int i = 0; i < 8; ++i) b[off + i] = (uint8_t)(v >> (8 * i));
}
uint64_t get_u64(const uint8_t *b, size_t off)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | b[off + i];
    return v;
}

bool mul_ok(uint64_t a, uint64_t b, uint64_t *out)
{
    if (a != 0 && b > UINT64_MAX / a) return false;
    *out = a * b;
    return true;
}

// valid_dimensions / valid_extents (mpi_io.rs:1931-1942)
bool valid_dimensions(const std::vector<uint64_t> &v)
{
    if (v.empty()) return false;
    uint64_t p = 1;
    for (uint64_t x : v)
        if (!mul_ok(p, x, &p)) return false;
    return true;
}
bool valid_extents(const std::vector<uint64_t> &v)
{
    if (!valid_dimensions(v)) return false;
    for (uint64_t x : v)
        if (x == 0) return false;
    return true;
}

struct Header {
    uint64_t n = 0, type = 0, width = 0, header_len = 0, payload_offset = 0, payload_bytes = 0;
    std::vector<uint64_t> global, extra, grid, perm;
};

// Header::parse_prefix (mpi_io.rs:1610-1631) with V1Fields checks (:1761-1825)
pfh_status parse_prefix(const uint8_t *p, Header &h, uint64_t *extra_rank, uint64_t *grid_rank)
{
    if (std::memcmp(p, kMagic, 8) != 0) return err(PFH_ERR_IO_INVALID_FILE, "magic");
    if (get_u64(p, 8) != kVersion) return err(PFH_ERR_IO_INVALID_FILE, "version");
    const uint64_t marker = get_u64(p, 24);
    if (marker == kIncomplete) return err(PFH_ERR_IO_INCOMPLETE_FILE, "file has no completed commit marker");
    if (marker != kCommit) return err(PFH_ERR_IO_INVALID_FILE, "commit marker");
    h.header_len = get_u64(p, 16);
    h.n = get_u64(p, 32);
    h.type = get_u64(p, 40);
    h.width = get_u64(p, 48);
    *extra_rank = get_u64(p, 56);
    h.payload_offset = get_u64(p, 64);
    h.payload_bytes = get_u64(p, 72);
    *grid_rank = get_u64(p, 80);
    if (h.n > kMaxRank || *extra_rank > kMaxRank || h.n + *extra_rank > kMaxRank || *grid_rank > kMaxRank ||
        *grid_rank > h.n)
        return err(PFH_ERR_IO_INVALID_FILE, "rank bound");
    const uint64_t expected = kPrefix + 8 * (h.n + *extra_rank + *grid_rank + h.n);
    if (h.header_len != expected || h.header_len > kMaxHeader) return err(PFH_ERR_IO_INVALID_FILE, "header length");
    if (h.payload_offset != h.header_len) return err(PFH_ERR_IO_INVALID_FILE, "payload offset");
    if (type_width(h.type) == 0 || (uint64_t)type_width(h.type) != h.width)
        return err(PFH_ERR_IO_INVALID_FILE, "type descriptor");
    return PFH_OK;
}

// Header::decode (mpi_io.rs:1633-1690): the arrays and their consistency
pfh_status decode_arrays(const uint8_t *b, uint64_t extra_rank, uint64_t grid_rank, Header &h)
{
    if (get_u64(b, 88) != 0) return err(PFH_ERR_IO_INVALID_FILE, "reserved field");
    size_t off = kPrefix;
    auto take = [&](uint64_t count, std::vector<uint64_t> &out) {
        out.resize(count);
        for (uint64_t i = 0; i < count; ++i, off += 8) out[i] = get_u64(b, off);
    };
    take(h.n, h.global);
    take(extra_rank, h.extra);
    take(grid_rank, h.grid);
    take(h.n, h.perm);
    // check_writer_grid: a product overflow and a non-positive extent are both
    // "header metadata"; so is a grid longer than the spatial rank
    std::vector<bool> seen(h.n, false);
    bool perm_ok = h.perm.size() == h.n;
    for (uint64_t v : h.perm) {
        if (v >= h.n || seen[v]) {
            perm_ok = false;
            break;
        }
        seen[v] = true;
    }
    if (!valid_dimensions(h.global) || !valid_extents(h.grid) || h.grid.size() > h.global.size() || !perm_ok)
        return err(PFH_ERR_IO_INVALID_FILE, "header metadata");
    uint64_t elems = 1;
    for (uint64_t v : h.global)
        if (!mul_ok(elems, v, &elems)) return err(PFH_ERR_IO_INVALID_FILE, "payload element count");
    for (uint64_t v : h.extra)
        if (!mul_ok(elems, v, &elems)) return err(PFH_ERR_IO_INVALID_FILE, "payload element count");
    uint64_t bytes = 0;
    if (!mul_ok(elems, h.width, &bytes) || bytes != h.payload_bytes) return err(PFH_ERR_IO_INVALID_FILE, "payload bytes");
    return PFH_OK;
}

struct Fd {
    int fd = -1;
    ~Fd()
    {
        if (fd >= 0) ::close(fd);
    }
};

bool read_all(int fd, void *buf, size_t n, uint64_t off)
{
    uint8_t *p = static_cast<uint8_t *>(buf);
    while (n > 0) {
        const ssize_t r = ::pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

bool write_all(int fd, const void *buf, size_t n, uint64_t off)
{
    const uint8_t *p = static_cast<const uint8_t *>(buf);
    while (n > 0) {
        const ssize_t r = ::pwrite(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

// Two pinned transfer chunks, freed on scope exit.
struct Pinned {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Pinned()
    {
        for (int i = 0; i < 2; ++i) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    hipError_t init(size_t bytes)
    {
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipHostMalloc(&buf[i], bytes, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

struct ArrayInfo {
    int n = 0, extra_rank = 0, width = 0;
    uint64_t global[PFH_MAX_DIMS] = {0, 0, 0, 0};
    uint64_t extra[PFH_MAX_DIMS] = {0, 0, 0, 0};
    int perm[PFH_MAX_DIMS] = {0, 1, 2, 3};
    uint64_t spatial = 1, extra_count = 1, elements = 1, bytes = 0;
    bool identity = true;
};

pfh_status array_info(const pfh_array *a, ArrayInfo &I)
{
    if (!a) return err(PFH_ERR_IO_INVALID_INPUT, "array");
    const pfh_pencil_layout &p = a->pencil;
    if (p.ndim < 1 || p.ndim > PFH_MAX_DIMS || a->extra_ndim < 0 || a->extra_ndim > PFH_MAX_DIMS ||
        p.topology_dims < 0 || p.topology_dims > p.ndim)
        return err(PFH_ERR_IO_INVALID_INPUT, "rank");
    I.width = type_width((uint64_t)a->scalar_type);
    if (I.width == 0) return err(PFH_ERR_IO_INVALID_INPUT, "scalar type");
    I.n = p.ndim;
    I.extra_rank = a->extra_ndim;
    bool seen[PFH_MAX_DIMS] = {false, false, false, false};
    for (int d = 0; d < I.n; ++d) {
        if (p.perm[d] < 0 || p.perm[d] >= I.n || seen[p.perm[d]]) return err(PFH_ERR_INVALID_PERMUTATION, "perm");
        seen[p.perm[d]] = true;
        I.perm[d] = p.perm[d];
        I.identity = I.identity && p.perm[d] == d;
        I.global[d] = p.shape[d];
        if (!mul_ok(I.spatial, p.shape[d], &I.spatial)) return err(PFH_ERR_IO_SIZE_LIMIT, "element count");
    }
    for (int d = 0; d < I.extra_rank; ++d) {
        I.extra[d] = a->extra_shape[d];
        if (!mul_ok(I.extra_count, a->extra_shape[d], &I.extra_count))
            return err(PFH_ERR_IO_SIZE_LIMIT, "element count");
    }
    if (!mul_ok(I.spatial, I.extra_count, &I.elements) || !mul_ok(I.elements, (uint64_t)I.width, &I.bytes) ||
        I.bytes > (uint64_t)INT64_MAX)
        return err(PFH_ERR_IO_SIZE_LIMIT, "global payload bytes");
    if (I.bytes > 0 && !a->data) return err(PFH_ERR_IO_INVALID_INPUT, "array data");
    return PFH_OK;
}

pfh_status check_path(const char *path)
{
    if (!path || !*path) return err(PFH_ERR_IO_INVALID_PATH, "path cannot be represented as a native path");
    if (std::strlen(path) > kMaxDescriptor) return err(PFH_ERR_IO_SIZE_LIMIT, "descriptor");
    return PFH_OK;
}

// Permutes between the pencil's memory order (`to_canonical` false: source
// canonical, destination memory order) with 1- or 2-byte words, as the
// transpose plans do (bit exact for every element type).
hipError_t permute(const ArrayInfo &I, bool to_canonical, const void *src, void *dst, hipStream_t s)
{
    const int word = I.width == 1 ? 1 : 2;
    const int64_t wpe = I.width / word;
    int src_perm[PFH_MAX_DIMS], dst_perm[PFH_MAX_DIMS];
    for (int d = 0; d < I.n; ++d) {
        src_perm[d] = to_canonical ? I.perm[d] : d;
        dst_perm[d] = to_canonical ? d : I.perm[d];
    }
    // source stride (in elements) of each logical axis
    int64_t sst[PFH_MAX_DIMS];
    int64_t st = 1;
    for (int q = I.n - 1; q >= 0; --q) {
        sst[src_perm[q]] = st;
        st *= (int64_t)I.global[src_perm[q]];
    }
    PermuteArgs pa{};
    pa.nd = I.n + 1;
    pa.len = (int64_t)I.spatial * wpe;
    pa.extra = (int64_t)I.extra_count;
    for (int q = 0; q < I.n; ++q) {
        pa.dext[q] = (int64_t)I.global[dst_perm[q]];
        pa.sstride[q] = sst[dst_perm[q]] * wpe;
    }
    pa.dext[I.n] = wpe;
    pa.sstride[I.n] = 1;
    return launch_permute_words(pa, src, dst, word, s);
}

} // namespace

extern "C" {

pfh_status pfh_io_workspace_bytes(const pfh_array *array, size_t *bytes)
{
    ArrayInfo I;
    pfh_status st = array_info(array, I);
    if (st) return st;
    if (!bytes) return err(PFH_ERR_INVALID_ARGUMENT, "argument");
    *bytes = (size_t)I.bytes;
    return clear_error();
}

// The array's payload in canonical order at `file_offset`: permuted on the
// device (unless the pencil's order is canonical), then copied to the host
// in pinned chunks, each written while the next one is copied.
static pfh_status write_canonical(int fd, const ArrayInfo &I, const pfh_array *array, void *workspace,
                                  uint64_t file_offset, hipStream_t s)
{
    if (I.bytes == 0) return PFH_OK;
    Pinned pin;
    IO_HIP(pin.init(std::min<uint64_t>(I.bytes, kChunk)));
    const void *canon = array->data;
    if (!I.identity) {
        IO_HIP(permute(I, true, array->data, workspace, s));
        canon = workspace;
    }
    const uint64_t nchunks = (I.bytes + kChunk - 1) / kChunk;
    auto issue = [&](uint64_t k) -> hipError_t {
        const uint64_t o = k * kChunk, len = std::min<uint64_t>(kChunk, I.bytes - o);
        hipError_t e = hipMemcpyAsync(pin.buf[k & 1], static_cast<const uint8_t *>(canon) + o, len,
                                      hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipEventRecord(pin.ev[k & 1], s);
    };
    IO_HIP(issue(0));
    for (uint64_t k = 0; k < nchunks; ++k) {
        IO_HIP(hipEventSynchronize(pin.ev[k & 1]));
        if (k + 1 < nchunks) IO_HIP(issue(k + 1));
        const uint64_t o = k * kChunk, len = std::min<uint64_t>(kChunk, I.bytes - o);
        if (!write_all(fd, pin.buf[k & 1], (size_t)len, file_offset + o)) {
            (void)hipStreamSynchronize(s);
            return err(PFH_ERR_IO_WRITE_INCOMPLETE, "payload write");
        }
    }
    return PFH_OK;
}

pfh_status pfh_io_write(const char *path, const pfh_array *array, void *workspace, size_t workspace_bytes,
                        pfh_stream_t stream)
{
    pfh_status st = check_path(path);
    if (st) return st;
    ArrayInfo I;
    if ((st = array_info(array, I))) return st;
    const int M = array->pencil.topology_dims > 0 ? array->pencil.topology_dims : 1;
    const uint64_t header_len = kPrefix + 8 * (uint64_t)(I.n + I.extra_rank + M + I.n);
    if (!I.identity && I.bytes > 0 && (!workspace || workspace_bytes < I.bytes))
        return set_error(PFH_ERR_WORKSPACE_TOO_SMALL, "canonical staging", (size_t)I.bytes,
                         workspace ? workspace_bytes : 0, 0, 0, 0);
    // Header::for_write / encode_into (mpi_io.rs:1492-1607)
    std::vector<uint8_t> hdr(header_len, 0);
    std::memcpy(hdr.data(), kMagic, 8);
    put_u64(hdr.data(), 8, kVersion);
    put_u64(hdr.data(), 16, header_len);
    put_u64(hdr.data(), 24, kIncomplete);
    put_u64(hdr.data(), 32, (uint64_t)I.n);
    put_u64(hdr.data(), 40, (uint64_t)array->scalar_type);
    put_u64(hdr.data(), 48, (uint64_t)I.width);
    put_u64(hdr.data(), 56, (uint64_t)I.extra_rank);
    put_u64(hdr.data(), 64, header_len);
    put_u64(hdr.data(), 72, I.bytes);
    put_u64(hdr.data(), 80, (uint64_t)M);
    put_u64(hdr.data(), 88, 0);
    size_t off = kPrefix;
    for (int d = 0; d < I.n; ++d, off += 8) put_u64(hdr.data(), off, I.global[d]);
    for (int d = 0; d < I.extra_rank; ++d, off += 8) put_u64(hdr.data(), off, I.extra[d]);
    for (int d = 0; d < M; ++d, off += 8) put_u64(hdr.data(), off, 1); // one-rank writer grid
    for (int d = 0; d < I.n; ++d, off += 8) put_u64(hdr.data(), off, (uint64_t)I.perm[d]);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // exclusive creation: an existing file is never replaced
    Fd f;
    f.fd = ::open(path, O_WRONLY | O_CREAT | O_EXCL | O_CLOEXEC, 0644);
    if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
    if (!write_all(f.fd, hdr.data(), hdr.size(), 0)) return err(PFH_ERR_IO_WRITE_INCOMPLETE, "header write");
    if ((st = write_canonical(f.fd, I, array, workspace, header_len, s))) return st;
    if (::fsync(f.fd) != 0) return err(PFH_ERR_IO_WRITE_INCOMPLETE, "payload flush");
    // commit marker at byte 24, flushed (write_payload_and_commit, mpi_io.rs:1146-1230)
    uint8_t marker[8];
    put_u64(marker, 0, kCommit);
    if (!write_all(f.fd, marker, 8, 24)) return err(PFH_ERR_IO_COMMIT_UNCERTAIN, "commit marker write");
    if (::fsync(f.fd) != 0) return err(PFH_ERR_IO_COMMIT_UNCERTAIN, "commit marker flush");
    const int fd = f.fd;
    f.fd = -1;
    if (::close(fd) != 0) return set_error(PFH_ERR_IO_NATIVE, "close", 0, (size_t)errno, 0, 0, 0);
    return clear_error();
}

// Opens and validates a file's header (read_mpi_inner, mpi_io.rs:600-760).
static pfh_status open_checked(const char *path, Fd &f, Header &h, uint64_t *file_size)
{
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
    struct stat sb;
    if (::fstat(f.fd, &sb) != 0) return set_error(PFH_ERR_IO_NATIVE, "fstat", 0, (size_t)errno, 0, 0, 0);
    *file_size = (uint64_t)sb.st_size;
    if (*file_size < kPrefix) return err(PFH_ERR_IO_INVALID_FILE, "short header");
    uint8_t prefix[kPrefix];
    if (!read_all(f.fd, prefix, kPrefix, 0)) return set_error(PFH_ERR_IO_NATIVE, "read", 0, (size_t)errno, 0, 0, 0);
    uint64_t extra_rank = 0, grid_rank = 0;
    pfh_status st = parse_prefix(prefix, h, &extra_rank, &grid_rank);
    if (st) return st;
    if (h.header_len > *file_size) return err(PFH_ERR_IO_INVALID_FILE, "truncated header");
    std::vector<uint8_t> full(h.header_len);
    if (!read_all(f.fd, full.data(), full.size(), 0))
        return set_error(PFH_ERR_IO_NATIVE, "read", 0, (size_t)errno, 0, 0, 0);
    return decode_arrays(full.data(), extra_rank, grid_rank, h);
}

// Reads [offset, offset + I.bytes) of the open file into the device staging
// buffer through the pinned chunks, byte swapping components when asked.
static pfh_status stage_payload(int fd, uint64_t offset, const ArrayInfo &I, void *staging, bool swap, int comp,
                                hipStream_t s)
{
    Pinned pin;
    if (I.bytes == 0) return PFH_OK;
    IO_HIP(pin.init(std::min<uint64_t>(I.bytes, kChunk)));
    const uint64_t nchunks = (I.bytes + kChunk - 1) / kChunk;
    for (uint64_t k = 0; k < nchunks; ++k) {
        const uint64_t o = k * kChunk, len = std::min<uint64_t>(kChunk, I.bytes - o);
        if (k >= 2) IO_HIP(hipEventSynchronize(pin.ev[k & 1])); // the copy out of this buffer is done
        uint8_t *b = static_cast<uint8_t *>(pin.buf[k & 1]);
        if (!read_all(fd, b, (size_t)len, offset + o)) {
            (void)hipStreamSynchronize(s);
            return set_error(PFH_ERR_IO_NATIVE, "read", 0, (size_t)errno, 0, 0, 0);
        }
        if (swap && comp > 1)
            for (uint64_t i = 0; i + comp <= len; i += comp)
                for (int a = 0, z = comp - 1; a < z; ++a, --z) std::swap(b[i + a], b[i + z]);
        IO_HIP(hipMemcpyAsync(static_cast<uint8_t *>(staging) + o, b, len, hipMemcpyHostToDevice, s));
        IO_HIP(hipEventRecord(pin.ev[k & 1], s));
    }
    IO_HIP(hipStreamSynchronize(s));
    return PFH_OK;
}

static pfh_status commit_staged(const ArrayInfo &I, const pfh_array *dst, const void *staging, hipStream_t s)
{
    if (I.bytes == 0) return PFH_OK;
    if (I.identity) IO_HIP(hipMemcpyAsync(dst->data, staging, I.bytes, hipMemcpyDeviceToDevice, s));
    else IO_HIP(permute(I, false, staging, dst->data, s));
    return PFH_OK;
}

pfh_status pfh_io_read(const char *path, const pfh_array *array, void *workspace, size_t workspace_bytes,
                       pfh_stream_t stream)
{
    pfh_status st = check_path(path);
    if (st) return st;
    ArrayInfo I;
    if ((st = array_info(array, I))) return st;
    if (I.bytes > 0 && (!workspace || workspace_bytes < I.bytes))
        return set_error(PFH_ERR_WORKSPACE_TOO_SMALL, "read staging", (size_t)I.bytes, workspace ? workspace_bytes : 0,
                         0, 0, 0);
    Fd f;
    Header h;
    uint64_t size = 0;
    if ((st = open_checked(path, f, h, &size))) return st;
    // Header::matches_view (mpi_io.rs:1692-1748)
    if (h.n != (uint64_t)I.n || h.type != (uint64_t)array->scalar_type || h.width != (uint64_t)I.width)
        return err(PFH_ERR_IO_METADATA_MISMATCH, "type or rank");
    bool same = h.extra.size() == (size_t)I.extra_rank;
    for (int d = 0; same && d < I.extra_rank; ++d) same = h.extra[d] == I.extra[d];
    if (!same) return err(PFH_ERR_IO_METADATA_MISMATCH, "extra shape");
    for (int d = 0; d < I.n; ++d)
        if (h.global[d] != I.global[d]) return err(PFH_ERR_IO_METADATA_MISMATCH, "global shape");
    if (h.payload_bytes != I.bytes) return err(PFH_ERR_IO_METADATA_MISMATCH, "payload size");
    if (h.payload_offset > UINT64_MAX - h.payload_bytes) return err(PFH_ERR_IO_INVALID_FILE, "payload size overflow");
    if (h.payload_offset + h.payload_bytes != size) return err(PFH_ERR_IO_INVALID_FILE, "trailing or truncated payload");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = stage_payload(f.fd, h.payload_offset, I, workspace, false, 1, s))) return st;
    if ((st = commit_staged(I, array, workspace, s))) return st;
    return clear_error();
}

pfh_status pfh_io_read_raw(const char *path, const pfh_array *array, uint64_t offset, int byte_order,
                           void *workspace, size_t workspace_bytes, pfh_stream_t stream)
{
    pfh_status st = check_path(path);
    if (st) return st;
    if (byte_order < PFH_BYTE_ORDER_NATIVE || byte_order > PFH_BYTE_ORDER_BIG)
        return err(PFH_ERR_IO_INVALID_INPUT, "byte order");
    ArrayInfo I;
    if ((st = array_info(array, I))) return st;
    if (offset > UINT64_MAX - I.bytes || offset + I.bytes > (uint64_t)INT64_MAX)
        return err(PFH_ERR_IO_SIZE_LIMIT, "raw payload end");
    if (I.bytes > 0 && (!workspace || workspace_bytes < I.bytes))
        return set_error(PFH_ERR_WORKSPACE_TOO_SMALL, "read staging", (size_t)I.bytes, workspace ? workspace_bytes : 0,
                         0, 0, 0);
    Fd f;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
    struct stat sb;
    if (::fstat(f.fd, &sb) != 0) return set_error(PFH_ERR_IO_NATIVE, "fstat", 0, (size_t)errno, 0, 0, 0);
    if ((uint64_t)sb.st_size < offset + I.bytes) return err(PFH_ERR_IO_INVALID_FILE, "truncated raw payload");
    // every real component is decoded in the requested order (a complex value
    // is two components); the host is little endian
    const bool swap = byte_order == PFH_BYTE_ORDER_BIG;
    const bool complex_type = array->scalar_type == PFH_SCALAR_C64 || array->scalar_type == PFH_SCALAR_C128;
    const int comp = complex_type ? I.width / 2 : I.width;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = stage_payload(f.fd, offset, I, workspace, swap, comp, s))) return st;
    if ((st = commit_staged(I, array, workspace, s))) return st;
    return clear_error();
}

// ---------------------------------------------------------------------------
// Named (v2) append-only container (named_mpi.rs:1-1470): a 32-byte header
// ("PIONAM02", version 2, header size 32, reserved 0), then records of a
// 72-byte header ("PIOREC02", version, name length, metadata length, payload
// length, total length, type code, width, reserved 0), the UTF-8 name, the
// metadata (spatial rank, extra rank, spatial dims, extra dims), the
// canonical payload and a trailing commit word. The committed prefix is
// found by scanning; the first malformed or uncommitted record ends it.
// ---------------------------------------------------------------------------

constexpr size_t kContainer = 32;
constexpr size_t kRecordHeader = 72;
constexpr uint64_t kNamedVersion = 2;
constexpr size_t kMaxName = 1024;
const char kNamedMagic[8] = {'P', 'I', 'O', 'N', 'A', 'M', '0', '2'};
const char kRecordMagic[8] = {'P', 'I', 'O', 'R', 'E', 'C', '0', '2'};

// valid_name_bytes (lib.rs:343-348): non-empty, at most 1024 bytes, no NUL,
// well-formed UTF-8 (no overlong forms, surrogates or code points above
// U+10FFFF, as Rust's from_utf8 requires).
static bool valid_name(const char *name, size_t *len)
{
    if (!name) return false;
    const size_t n = std::strlen(name); // a NUL ends the C string: embedded NULs cannot be passed
    *len = n;
    if (n == 0 || n > kMaxName) return false;
    const auto *b = reinterpret_cast<const uint8_t *>(name);
    for (size_t i = 0; i < n;) {
        const uint8_t c = b[i];
        int k;
        uint32_t cp, lo;
        if (c < 0x80) {
            ++i;
            continue;
        } else if ((c & 0xE0) == 0xC0) {
            k = 1, cp = c & 0x1F, lo = 0x80;
        } else if ((c & 0xF0) == 0xE0) {
            k = 2, cp = c & 0x0F, lo = 0x800;
        } else if ((c & 0xF8) == 0xF0) {
            k = 3, cp = c & 0x07, lo = 0x10000;
        } else {
            return false;
        }
        for (int j = 1; j <= k; ++j) {
            if (i + j >= n || (b[i + j] & 0xC0) != 0x80) return false;
            cp = (cp << 6) | (b[i + j] & 0x3F);
        }
        if (cp < lo || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return false;
        i += k + 1;
    }
    return true;
}

struct NamedRecord {
    std::string name;
    uint64_t type = 0, width = 0, payload_offset = 0, payload = 0;
    std::vector<uint64_t> global, extra;
};

// scan_root (named_mpi.rs:333-505): the committed records and the end of the
// committed prefix.
static pfh_status scan_named(int fd, uint64_t size, std::vector<NamedRecord> &out, uint64_t *tail)
{
    uint8_t h[kContainer];
    if (size < kContainer || !read_all(fd, h, kContainer, 0) || std::memcmp(h, kNamedMagic, 8) != 0 ||
        get_u64(h, 8) != kNamedVersion || get_u64(h, 16) != kContainer || get_u64(h, 24) != 0)
        return err(PFH_ERR_IO_INVALID_FILE, "named container header");
    uint64_t p = kContainer;
    while (p < size) {
        if (size - p < kRecordHeader) break;
        uint8_t rh[kRecordHeader];
        if (!read_all(fd, rh, kRecordHeader, p)) return err(PFH_ERR_IO_INVALID_FILE, "named metadata read");
        if (std::memcmp(rh, kRecordMagic, 8) != 0 || get_u64(rh, 8) != kNamedVersion || get_u64(rh, 64) != 0) break;
        const uint64_t nl = get_u64(rh, 16), ml = get_u64(rh, 24), payload = get_u64(rh, 32), total = get_u64(rh, 40);
        if (nl == 0 || nl > kMaxName || ml > kMaxHeader) break;
        if (total > size - p || p + total > (uint64_t)INT64_MAX) break;
        if (payload > total || kRecordHeader + nl + ml + payload + 8 != total) break;
        std::vector<uint8_t> body(nl + ml);
        if (!read_all(fd, body.data(), body.size(), p + kRecordHeader))
            return err(PFH_ERR_IO_INVALID_FILE, "named metadata read");
        std::string name(reinterpret_cast<const char *>(body.data()), nl);
        size_t len = 0;
        if (std::memchr(name.data(), 0, nl) || !valid_name(name.c_str(), &len) || len != nl) break;
        bool dup = false;
        for (const NamedRecord &r : out) dup = dup || r.name == name;
        if (dup) break;
        if (ml < 16) break;
        const uint8_t *meta = body.data() + nl;
        const uint64_t nr = get_u64(meta, 0), er = get_u64(meta, 8);
        if (nr == 0 || nr > kMaxRank || er > kMaxRank || nr + er > kMaxRank || 16 + 8 * (nr + er) != ml) break;
        std::vector<uint64_t> dims(nr + er);
        uint64_t elements = 1;
        bool ok = true;
        for (uint64_t i = 0; i < nr + er && ok; ++i) {
            dims[i] = get_u64(meta, 16 + 8 * i);
            ok = mul_ok(elements, dims[i], &elements);
        }
        if (!ok) break;
        const uint64_t type = get_u64(rh, 48), width = get_u64(rh, 56);
        uint64_t bytes = 0;
        if (type_width(type) == 0 || (uint64_t)type_width(type) != width || !mul_ok(elements, width, &bytes) ||
            bytes != payload)
            break;
        uint8_t marker[8];
        if (!read_all(fd, marker, 8, p + total - 8)) return err(PFH_ERR_IO_INVALID_FILE, "named metadata read");
        if (get_u64(marker, 0) != kCommit) break;
        NamedRecord r;
        r.name = name;
        r.type = type;
        r.width = width;
        r.payload = payload;
        r.global.assign(dims.begin(), dims.begin() + (ptrdiff_t)nr);
        r.extra.assign(dims.begin() + (ptrdiff_t)nr, dims.end());
        r.payload_offset = p + kRecordHeader + nl + ml;
        out.push_back(std::move(r));
        p += total;
    }
    *tail = p;
    return PFH_OK;
}

// record() (named_mpi.rs:920-980): the record header, name and metadata.
static std::vector<uint8_t> record_head(const char *name, size_t nl, const ArrayInfo &I, int type, uint64_t *total)
{
    const uint64_t ml = 16 + 8 * (uint64_t)(I.n + I.extra_rank);
    *total = kRecordHeader + nl + ml + I.bytes + 8;
    std::vector<uint8_t> h(kRecordHeader + nl + ml, 0);
    std::memcpy(h.data(), kRecordMagic, 8);
    put_u64(h.data(), 8, kNamedVersion);
    put_u64(h.data(), 16, nl);
    put_u64(h.data(), 24, ml);
    put_u64(h.data(), 32, I.bytes);
    put_u64(h.data(), 40, *total);
    put_u64(h.data(), 48, (uint64_t)type);
    put_u64(h.data(), 56, (uint64_t)I.width);
    put_u64(h.data(), 64, 0);
    std::memcpy(h.data() + kRecordHeader, name, nl);
    uint8_t *meta = h.data() + kRecordHeader + nl;
    put_u64(meta, 0, (uint64_t)I.n);
    put_u64(meta, 8, (uint64_t)I.extra_rank);
    for (int d = 0; d < I.n; ++d) put_u64(meta, 16 + 8 * d, I.global[d]);
    for (int d = 0; d < I.extra_rank; ++d) put_u64(meta, 16 + 8 * (I.n + d), I.extra[d]);
    return h;
}

static pfh_status write_record(bool append, const char *path, const char *name, const pfh_array *array,
                               void *workspace, size_t workspace_bytes, pfh_stream_t stream)
{
    size_t nl = 0;
    if (!valid_name(name, &nl)) return err(PFH_ERR_IO_INVALID_NAME, "invalid dataset name");
    pfh_status st = check_path(path);
    if (st) return st;
    ArrayInfo I;
    if ((st = array_info(array, I))) return st;
    if (!I.identity && I.bytes > 0 && (!workspace || workspace_bytes < I.bytes))
        return set_error(PFH_ERR_WORKSPACE_TOO_SMALL, "canonical staging", (size_t)I.bytes,
                         workspace ? workspace_bytes : 0, 0, 0, 0);
    uint64_t total = 0;
    const std::vector<uint8_t> head = record_head(name, nl, I, array->scalar_type, &total);
    Fd f;
    uint64_t offset = kContainer;
    if (append) {
        f.fd = ::open(path, O_RDWR | O_CLOEXEC);
        if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
        struct stat sb;
        if (::fstat(f.fd, &sb) != 0) return set_error(PFH_ERR_IO_NATIVE, "fstat", 0, (size_t)errno, 0, 0, 0);
        std::vector<NamedRecord> recs;
        uint64_t tail = 0;
        if ((st = scan_named(f.fd, (uint64_t)sb.st_size, recs, &tail))) return st;
        if (tail != (uint64_t)sb.st_size) return err(PFH_ERR_IO_INVALID_TAIL, "dataset container has an invalid tail");
        for (const NamedRecord &r : recs)
            if (r.name == std::string(name, nl)) return err(PFH_ERR_IO_DUPLICATE_NAME, "dataset name already exists");
        offset = tail;
    } else {
        f.fd = ::open(path, O_RDWR | O_CREAT | O_EXCL | O_CLOEXEC, 0644);
        if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
        uint8_t ch[kContainer] = {0};
        std::memcpy(ch, kNamedMagic, 8);
        put_u64(ch, 8, kNamedVersion);
        put_u64(ch, 16, kContainer);
        if (!write_all(f.fd, ch, kContainer, 0)) return err(PFH_ERR_IO_WRITE_INCOMPLETE, "container header write");
    }
    if (offset > (uint64_t)INT64_MAX - total) return err(PFH_ERR_IO_SIZE_LIMIT, "named container");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!write_all(f.fd, head.data(), head.size(), offset)) return err(PFH_ERR_IO_WRITE_INCOMPLETE, "record header write");
    if ((st = write_canonical(f.fd, I, array, workspace, offset + head.size(), s))) return st;
    if (::fsync(f.fd) != 0) return err(PFH_ERR_IO_WRITE_INCOMPLETE, "record flush");
    uint8_t marker[8];
    put_u64(marker, 0, kCommit);
    if (!write_all(f.fd, marker, 8, offset + total - 8)) return err(PFH_ERR_IO_COMMIT_UNCERTAIN, "commit word write");
    if (::fsync(f.fd) != 0) return err(PFH_ERR_IO_COMMIT_UNCERTAIN, "commit word flush");
    const int fd = f.fd;
    f.fd = -1;
    if (::close(fd) != 0) return set_error(PFH_ERR_IO_NATIVE, "close", 0, (size_t)errno, 0, 0, 0);
    return clear_error();
}

pfh_status pfh_io_write_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                              size_t workspace_bytes, pfh_stream_t stream)
{
    return write_record(false, path, name, array, workspace, workspace_bytes, stream);
}

pfh_status pfh_io_append_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                               size_t workspace_bytes, pfh_stream_t stream)
{
    return write_record(true, path, name, array, workspace, workspace_bytes, stream);
}

pfh_status pfh_io_read_named(const char *path, const char *name, const pfh_array *array, void *workspace,
                             size_t workspace_bytes, pfh_stream_t stream)
{
    size_t nl = 0;
    if (!valid_name(name, &nl)) return err(PFH_ERR_IO_INVALID_NAME, "invalid dataset name");
    pfh_status st = check_path(path);
    if (st) return st;
    ArrayInfo I;
    if ((st = array_info(array, I))) return st;
    if (I.bytes > 0 && (!workspace || workspace_bytes < I.bytes))
        return set_error(PFH_ERR_WORKSPACE_TOO_SMALL, "read staging", (size_t)I.bytes, workspace ? workspace_bytes : 0,
                         0, 0, 0);
    Fd f;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return set_error(PFH_ERR_IO_NATIVE, "open", 0, (size_t)errno, 0, 0, 0);
    struct stat sb;
    if (::fstat(f.fd, &sb) != 0) return set_error(PFH_ERR_IO_NATIVE, "fstat", 0, (size_t)errno, 0, 0, 0);
    std::vector<NamedRecord> recs;
    uint64_t tail = 0;
    if ((st = scan_named(f.fd, (uint64_t)sb.st_size, recs, &tail))) return st;
    const NamedRecord *r = nullptr;
    for (const NamedRecord &x : recs)
        if (x.name == std::string(name, nl)) r = &x;
    if (!r) return err(PFH_ERR_IO_NOT_FOUND, "dataset name was not found");
    bool same = r->type == (uint64_t)array->scalar_type && r->width == (uint64_t)I.width &&
                r->global.size() == (size_t)I.n && r->extra.size() == (size_t)I.extra_rank;
    for (int d = 0; same && d < I.n; ++d) same = r->global[d] == I.global[d];
    for (int d = 0; same && d < I.extra_rank; ++d) same = r->extra[d] == I.extra[d];
    if (!same) return err(PFH_ERR_IO_METADATA_MISMATCH, "named record");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = stage_payload(f.fd, r->payload_offset, I, workspace, false, 1, s))) return st;
    if ((st = commit_staged(I, array, workspace, s))) return st;
    return clear_error();
}

pfh_status pfh_io_catalog(const char *path, pfh_dataset_info *info)
{
    pfh_status st = check_path(path);
    if (st) return st;
    if (!info) return err(PFH_ERR_INVALID_ARGUMENT, "info");
    Fd f;
    Header h;
    uint64_t size = 0;
    if ((st = open_checked(path, f, h, &size))) return st;
    if (h.n > PFH_MAX_DIMS || h.extra.size() > PFH_MAX_DIMS || h.grid.size() > PFH_MAX_DIMS)
        return err(PFH_ERR_UNSUPPORTED, "more than 4 dimensions");
    pfh_dataset_info out;
    std::memset(&out, 0, sizeof(out));
    out.scalar_type = (int)h.type;
    out.ndim = (int)h.n;
    out.extra_ndim = (int)h.extra.size();
    out.grid_ndim = (int)h.grid.size();
    for (size_t d = 0; d < h.n; ++d) {
        out.shape[d] = h.global[d];
        out.writer_perm[d] = (int)h.perm[d];
    }
    for (size_t d = 0; d < h.extra.size(); ++d) out.extra_shape[d] = h.extra[d];
    for (size_t d = 0; d < h.grid.size(); ++d) out.writer_grid[d] = h.grid[d];
    out.payload_bytes = h.payload_bytes;
    *info = out;
    return clear_error();
}

} // extern "C"

// Host-side planning structures of the gfx950 FFT path.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pencil_fft_hip.h"
#include "pfh_kernels.hpp"

namespace pfh {

// A memory layout: logical shape, memory order (perm[pos] = logical axis at
// memory position pos, slowest first), extra blocks and element size. This
// is the part of a Pencil (crates/pencil-array/src/pencil.rs:69-112) that a
// one-rank process grid needs.
struct Layout {
    int N = 0;
    std::array<int, 8> perm{};
    std::array<int64_t, 8> shape{};
    int64_t extra = 1;
    int elem = 0; // bytes

    int64_t len() const
    {
        int64_t p = 1;
        for (int a = 0; a < N; ++a) p *= shape[a];
        return p;
    }
    int64_t total() const { return len() * extra; }
    int64_t bytes() const { return total() * elem; }
    // element stride of every logical axis
    std::array<int64_t, 8> strides() const
    {
        std::array<int64_t, 8> st{};
        int64_t run = 1;
        for (int p = N - 1; p >= 0; --p) {
            st[perm[p]] = run;
            run *= shape[perm[p]];
        }
        return st;
    }
    bool same_order(const Layout &o) const
    {
        if (N != o.N) return false;
        for (int p = 0; p < N; ++p)
            if (perm[p] != o.perm[p]) return false;
        return true;
    }
};

// BUF_X0..BUF_X2: workspace regions of the pieces of long lines.
enum BufId : int { BUF_SRC = 0, BUF_DST = 1, BUF_W0 = 2, BUF_W1 = 3, BUF_X0 = 4, BUF_X1 = 5, BUF_X2 = 6, BUF_COUNT = 7 };

// One kernel launch of a plan operation.
struct Step {
    int mode = M_COPY_C;
    int flags = 0;
    int n = 0, m = 0, nin = 0, nout = 0;
    std::vector<int> radices;
    double scale = 1.0, prew0 = 1.0, prew1 = 1.0, postw0 = 1.0, postw1 = 1.0;
    int axis = -1;
    bool in_cplx = true, out_cplx = true;
    Layout in, out;
    int in_buf = BUF_SRC, out_buf = BUF_DST;
    // launch shape
    LineGeom g{};
    int64_t nlines = 0, ips = 1, ops = 1;
    int in_col = 0, out_col = 0;
    int lt = 1, tpl = 1, ls = 1, threads = 64;
    int lc = 1; // lines computed at a time: generic stages may hold lt = lc * rounds lines
    int64_t tiles = 0, grid = 0;
    int xcd_per = 0;
    size_t lds_bytes = 0;
    // tables (complex entries, offsets relative to tw_base)
    int64_t tw_base = 0;
    int twn = 0;
    std::array<int, kMaxPasses> twoff{};
    int pre0 = 0, pre1 = 0, post0 = 0, post1 = 0;
    // R2C boundary acceptance check before this step
    bool endpoint_check = false;
    EndpointArgs ep{};
    bool local_endpoint_check = false;
    bool gated = false; // the check ran as its own launch; the stage still reads the gate
    // compile-time specialized kernel (k_line_fast) when fast is set
    bool fast = false;
    int fr[3] = {0, 0, 0};
    int fem = 16;
    int digit_axis[kMaxDigits] = {-1, -1, -1, -1, -1}; // logical axis of each line digit (-1: extras)
    // Long lines: a stage whose line does not fit one workgroup (or whose
    // FFT length has a large prime factor) runs as `pieces`, in order; a
    // piece is a generic line stage (launch 0) or the prep / post kernel of
    // a long real line (launch 1 / 2, long_kind a LongKind).
    bool long_only = false; // no single-workgroup radix plan: pieces required
    std::vector<Step> pieces;
    int launch = 0;
    int long_kind = 0;
    int64_t xtab_in = -1, xtab_out = -1, xtw = -1; // table offsets (complex entries), -1: none
    int xdig = -1, xmul_in = 1, xmul_out = 1, xvalid_in = 0, xvalid_out = 0, xtw_m = 0, xtw_bsh = 0;
};

// Two consecutive steps run slab by slab over an axis that neither
// transforms and that is the slowest memory axis of the first step's input
// and of the intermediate array: the intermediate of one slab lives in a
// small workspace region (T) that stays in the device's last-level cache
// between the two launches, instead of a full array in HBM.
struct SlabPair {
    int e = -1;                   // first step index (the pair is steps e, e+1)
    int axis = -1;                // slab axis
    int nslab = 0;                // number of full slabs
    int64_t extent = 0;           // planes of a full slab
    int64_t tail = 0;             // planes of the last, partial slab (0: none)
    Step first, second;           // slab-shaped steps of a full slab (first writes T, second reads T)
    Step first_tail, second_tail; // the same for the partial slab
    int64_t in_plane_bytes = 0;   // source advance per plane of the slab axis (first step's input)
    int64_t out_plane_bytes = 0;  // destination advance per plane (second step's output)
};

struct OpPlan {
    bool valid = false;
    std::vector<Step> steps;
    Layout in, out;          // layouts of src and dst for this op
    int64_t w_bytes[2] = {0, 0};
    std::vector<SlabPair> pairs; // slab-pipelined step pairs (disjoint, ascending e)
    int64_t t_bytes = 0;         // slab intermediate region of the workspace
    int64_t x_bytes[3] = {0, 0, 0}; // regions of the pieces of long lines (BUF_X0..BUF_X2)
    int64_t long_lines = 0;      // local plans: lines per chunk of a long-line operation (0: no pieces)
};

struct DeviceInfo {
    int device = 0;
    size_t lds_limit = 65536;
    int cus = 256;
    char arch[64] = {}; // architecture name with its feature suffixes (gfx950:sramecc+:xnack-)
};

// Planner knobs (defaults, or chosen by the batch autotuner).
struct Tuning {
    int col_lines = 0;   // lines per tile on column-access stages (0: 128 bytes per row)
    int max_radix = 0;   // largest radix of the generic kernel (0: 16 for f32, 8 for f64)
    int block = 0;       // target threads per block (0: 256)
    int block_cap = 0;   // upper bound on threads per block (0: the kernel's launch bound)
    bool allow_fast = true;
    bool allow_xcd = true;
    int slab_mb = -1;    // slab intermediate target in MiB (0: no slab pipelining, -1: default)
    bool fill_block = false; // column-access tiles also take `block` threads (batched launches)
    bool batch_multi = true; // batched stages: one launch over every compile-time plan
    bool batch_persistent = true; // batched launches: persistent workgroups (resident count) looping
                                  // over the tiles, else one workgroup per tile
    int batch_group_mb = 0;  // batched problems run group by group (all stages of a group before
                             // the next; groups share one workspace region); 0: one group
    int batch_streams = 1;   // batched launches: 2 runs the problems whose first pass is a plane stage
                             // on a second stream, concurrently with the others (no dependency between them)
    bool batch_plane = true; // batched 3D f32 problems: the two inner route stages fused into one
                             // plane stage where both lengths have a plane plan and a plane fits in LDS
};

struct PlanImpl {
    pfh_desc desc{};
    int precision = PFH_F32;
    int kind = 0;
    int N = 0;
    int64_t extra = 1;
    int64_t partials = 1; // endpoint-statistics chunks kept in the workspace status area
    uint64_t tag = 0;     // unique per plan: a workspace's status area records the plan it is bound to
    Tuning tuning;
    DeviceInfo dev;
    // op -> plan (out of place and in place)
    OpPlan ops[6];
    OpPlan ops_inplace[6];
    // twiddle/pre/post tables of every step of every op, one device buffer
    std::vector<double> host_tables; // complex pairs, f64
    std::map<std::string, int64_t> table_index; // long-line tables already in host_tables (name -> offset)
    bool tables_frozen = false;                 // host_tables uploaded: no table may be added
    void *dev_tables = nullptr;
    int64_t table_entries = 0;
    // transposes
    Layout tr_src, tr_dst;
    int word = 0, words_per_elem = 1;
    // local plans
    size_t line_len = 0;
    int r2r_kind = 0;
    // local long-line plans: the expanded operation for a chunk of cnt
    // lines, keyed (op, cnt), built on the first execute that runs that
    // chunk size (executes may come from several host threads)
    struct ChunkCache {
        std::mutex mu;
        std::map<std::pair<int, int64_t>, std::shared_ptr<const OpPlan>> plans;
    };
    std::shared_ptr<ChunkCache> chunks = std::make_shared<ChunkCache>();
    ~PlanImpl();
};

} // namespace pfh

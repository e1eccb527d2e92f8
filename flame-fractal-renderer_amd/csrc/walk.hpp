// Walker kernel arguments and launch entry (host and device).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace ffr {

// device-side render statistics (buffer_renderer.hpp:85-106)
struct DevStats {
    unsigned long long s_iter;
    unsigned long long s_plot;
    unsigned long long n_bad;   // bad values seen; > limit stops new walkers
    unsigned long long atomics; // global atomic adds on the buffer (misses, flushes)
    unsigned long long flush_atomics; // of those, the LDS cell-cache flushes
    double pmin[3];
    double pmax[3];
};

constexpr int kMaxSlots = 64;     // variation slots per flame
constexpr int kMaxFinalVars = 16; // variations of the final xform
constexpr int kMaxXforms = 1024;  // xforms after removing weight 0

// The walker kernel's arguments in two parts.  WalkDyn holds what changes
// from launch to launch: device addresses, seed and walker positions, the
// bad-value limits, the LDS cell cache's geometry and the binned passes.
// WalkArgs adds what the flame fixes: bounds and cell geometry, the xform
// count, the table layout and the variation slots.  A flame compiled into
// its own kernel (rtc.cpp) holds the second part as compile-time constants
// of a type derived from WalkDyn; the kernel templates read both parts
// through `a.field` either way.
template <typename R> struct WalkDyn {
    const R *table;            // flame table (see DeviceFlame), copied to LDS
    const unsigned long long *masks; // [nxf] used-slot masks, copied to LDS
    void *buf;                 // cells * (1 + r) words
    DevStats *stats;
    unsigned long long *xf_count; // [nxf] selections by sorted position
    unsigned long long *bv_xfs;   // [bv_cap] sorted position of bad-value xforms
    double *bv_pts;               // [bv_cap][3]
    // hot-cell hints of the flame's LDS cell cache: [tag, key of slot 0,
    // key of slot 1, ...]; each task seeds its cache from them and leaves
    // its own hot cells for the next (nullptr: no hints)
    unsigned *hot;
    uint64_t seed;
    uint64_t walker_base;      // stream index of walker 0 of this launch
    uint64_t n_walkers;
    uint64_t batch;            // samples per walker
    uint64_t last_samples;     // samples of the last walker
    uint64_t bv_limit;
    uint64_t bv_cap;
    uint64_t task_walkers;     // walkers per workgroup task (one flame's launch)
    int settle;
    // LDS cell cache (kernels built with it): cache_slots entries of
    // [key u32 | count u32 | r colour words], probed cache_ways deep; when
    // cache_direct the cell index is the slot (the whole buffer fits)
    int cache_slots, cache_log2, cache_ways, cache_direct;
    // hints apply when hot[0] == hot_tag (the same cache geometry); a slot
    // is left as a hint when its task counted at least hot_thresh there
    unsigned hot_tag, hot_thresh;
    // binned accumulation (walk_bin_kernel): this pass runs iterations
    // [it0, it1) of every walker and writes each plotted sample's cell
    // index to rec (task-major, then iteration, then lane); walkers keep
    // their point and stream position in wstate between passes
    unsigned *rec;
    void *wstate;
    uint64_t it0, it1;
    // colour flames: the plotted samples' colours, component i of record k
    // at (R *)pay + i * pay_stride + k
    void *pay;
    uint64_t pay_stride;
};

template <typename R> struct WalkArgs : WalkDyn<R> {
    // the flame fields below are run-time values (a compiled flame's type
    // sets this true and holds them, and cw[], as constants)
    static constexpr bool kCompiled = false;
    R lo[3], hi[3], mult_d[3];
    uint64_t mult_i[3];
    uint64_t cells;
    int nxf, n_slots, has_final, fin_nv, r;
    int tab_reals;
    int xf0, fin0;
    // slot and final-variation descriptors live in the kernel arguments so
    // the kernel reads them with scalar loads
    int slot_type[kMaxSlots];
    int slot_off[kMaxSlots];    // table offset of the slot's entries
    int slot_stride[kMaxSlots]; // entry size
    int fin_type[kMaxFinalVars];
    int fin_off[kMaxFinalVars];
    // per-xform variation lists of all-basic 2-D flames (DeviceFlame::bt0);
    // bt_union: bit k set when some xform's list holds basic code k
    int bt0, bt_stride, bt_nvmax, bt_union;
    // xform record layout (flame_model.hpp XformLayout)
    int l_preA, l_preb, l_postA, l_postb, l_speed, l_color, l_has_color, l_has_pre,
        l_has_post, l_stride;
};

// The xform record layout (flame_model.cpp xform_layout): every offset but
// the stride depends only on the dimension count, so the kernels take them
// as constants (immediate offsets of the LDS reads, no address arithmetic)
template <int D> struct XformLay {
    static constexpr int preA = 0, preb = D * D, postA = D * D + D, postb = 2 * D * D + D,
                         speed = 2 * D * D + 2 * D, has_color = speed + 1, has_pre = speed + 2,
                         has_post = speed + 3, color = speed + 4;
};

// lanes per workgroup (= walkers per task) of the binned walker kernel
constexpr int kBinWalkBlock = 256;

// a walker between two passes of binned accumulation
template <typename R> struct WalkerState {
    R p[3];
    R c[4];     // colour (colour flames)
    uint64_t n; // stream position, or kStoppedWalker
};
constexpr uint64_t kStoppedWalker = ~0ull;

// one renderer's statistics, for the reset of a batched launch
struct ResetEntry {
    DevStats *st;
    unsigned long long *xf;
    int nxf;
    int pad;
};

// one task of a batched launch: walkers [w_begin, w_end) of flame `flame`
struct BatchTask {
    unsigned flame;
    unsigned pad;
    uint64_t w_begin, w_end;
};

// hot-cell hint slots kept per renderer (caches with more slots take none)
constexpr int kHotSlots = 1 << 15;

// the wave-per-walker kernel (chain_kernel.inc): variation slots per xform
// it holds, walkers (waves) per workgroup, and the most walkers per CU it
// takes (beyond, the lane-per-walker kernel has enough waves in flight).
// The limit depends on the length of its serial step.  Measured at 8 to 128
// walkers of 16384 per CU against the lane kernel (kernel ms, FP64): with
// identity post-affines and no colour it wins up to 64 (Sierpinski 19.9
// against 25.3, uneven weights 30.3 against 32.9) and breaks even at 128
// (39.3 against 40.7, 59.9 against 59.6); with post-affines from the table
// it wins at 16 (three linear slots 12.0 against 22.7) and loses at 32
// (23.6 against 22.8); colour builds win at 16 (61.2 against 63.8, with
// post-affines 38.4 against 39.1) and break even at 32.
constexpr int kChainSlots = 4;
constexpr int kChainWaves = 4;
constexpr int kChainWalkersPerCuPostId = 64; // identity post-affines, no colour
constexpr int kChainWalkersPerCuGeneral = 16; // post-affines from the table, or colour
__host__ __device__ constexpr int chain_walkers_per_cu(bool post_id, int r)
{
    return (r == 0 && post_id) ? kChainWalkersPerCuPostId : kChainWalkersPerCuGeneral;
}

// The wave-per-walker kernel stages each block's 64 steps in LDS, one row
// of reals per step holding what the serial part reads: the pre-affine
// offset; the matrix and slot weights unless one shared matrix and weight
// serve every xform (sh); the post-affine unless every one is the identity
// (pi); colour speed and colour (nc > 0); the flags where they select
// (3-D, slots, colour).  Rows are whole 16-byte units.
__host__ __device__ constexpr int chain_row_reals(int real_bytes, int d, int nc, int ns, bool pi,
                                                  bool sh)
{
    const int n = d + (sh ? 0 : d * d + ns) + (pi ? 0 : d * d + d) + (nc > 0 ? 1 + nc : 0) +
                  ((d == 3 || !sh || nc > 0) ? 1 : 0);
    const int per16 = 16 / real_bytes;
    return (n + per16 - 1) / per16 * per16;
}
// rows per wave: the block's 64 steps and two more that the serial part
// reads ahead past the last step (never used)
constexpr int kChainRows = 66;
// the kernel's build for a flame: slots rounded up to 1, 2 or 4; the
// shared-matrix build takes one slot and identity post-affines
__host__ __device__ constexpr int chain_build_ns(int slots)
{
    return slots <= 1 ? 1 : slots == 2 ? 2 : 4;
}
__host__ __device__ constexpr bool chain_build_sh(bool shared, bool post_id, int slots)
{
    return shared && post_id && slots <= 1;
}

#ifndef __HIPCC_RTC__
// a walker-kernel launch: one flame (args, a device WalkArgs) or a batch
// (descs, device WalkArgs[] indexed by the tasks' flame numbers)
struct WalkLaunch {
    int nc = 0, vs = 0;
    bool cache = false;
    int grid = 0, block = 0;
    size_t lds = 0;
    hipStream_t stream = nullptr;
    const void *args = nullptr;
    const void *descs = nullptr;
    const BatchTask *tasks = nullptr;
    const unsigned *ntasks = nullptr; // device word: the number of tasks
    unsigned *queue = nullptr;        // device word: next task (zero at launch)
    bool prepare_only = false;        // set the kernel's LDS limit, launch nothing
    bool chain = false;               // one flame, one walker per wave (chain_kernel)
    int chain_slots = 0;              // its variation slots (1..kChainSlots)
    bool chain_post_id = false;       // every xform's post-affine is the identity
    bool chain_shared = false;        // one pre-affine matrix and weight for all
    bool chain_diag = false;          // ... and that matrix diagonal
    bool bin = false;                 // binned accumulation (walk_bin_kernel)
};
#endif // __HIPCC_RTC__

constexpr unsigned kEmptyKey = 0xFFFFFFFFu;

// the basic variation set: linear, sinusoidal, spherical, swirl, horseshoe,
// polar (type ids 0, 1, 2, 20, 21, 22) and their codes 0..5
constexpr int kBasicTypes = 6;
__host__ __device__ constexpr int basic_type(int code) { return code < 3 ? code : code + 17; }
__host__ __device__ constexpr int basic_code(int type) { return type < 3 ? type : type - 17; }

// largest workgroup of the kernels that accumulate through the LDS cell
// cache: 1024 lanes (16 waves share one cache) where the kernel fits 128
// VGPRs, 512 for the FP64 and all-variation kernels, which need more
__host__ __device__ constexpr int cache_block_max(bool fp64, int vs)
{
    return (fp64 || vs == 2) ? 512 : 1024;
}

// LDS layout of the walker kernel: the flame table, then (8-byte aligned)
// the used-slot mask of each xform (u64), the per-xform selection counters
// (u64), then the cell cache
__host__ __device__ inline size_t lds_masks_offset(int tab_reals, int real_bytes)
{
    return ((size_t)tab_reals * (size_t)real_bytes + 7u) & ~(size_t)7u;
}

__host__ __device__ inline size_t cache_lds_offset(int tab_reals, int real_bytes, int nxf)
{
    return lds_masks_offset(tab_reals, real_bytes) + 16u * (size_t)nxf;
}

// the wave-per-walker kernel's step rows follow the table, masks and
// counters (16-byte aligned): kChainWaves waves of kChainRows rows
__host__ __device__ inline size_t chain_stage_offset(int tab_reals, int real_bytes, int nxf)
{
    return (cache_lds_offset(tab_reals, real_bytes, nxf) + 15u) & ~(size_t)15u;
}

// bytes of one cell-cache entry: key, count and the colour words
__host__ __device__ inline size_t cache_slot_bytes(int r, int colour_bytes)
{
    return 8u + (size_t)r * (size_t)colour_bytes;
}

// precision modes (num_t / hist_t / colour word)
enum { MODE_F64_U64 = 0, MODE_F32_U32 = 1, MODE_F32_U64 = 2 };

} // namespace ffr

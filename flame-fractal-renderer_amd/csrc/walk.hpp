// Walker kernel arguments and launch entry (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ffr {

// device-side render statistics (buffer_renderer.hpp:85-106)
struct DevStats {
    unsigned long long s_iter;
    unsigned long long s_plot;
    unsigned long long n_bad;   // bad values seen; > limit stops new walkers
    unsigned long long pad;
    double pmin[3];
    double pmax[3];
};

constexpr int kMaxSlots = 64;     // variation slots per flame
constexpr int kMaxFinalVars = 16; // variations of the final xform
constexpr int kMaxXforms = 1024;  // xforms after removing weight 0

template <typename R> struct WalkArgs {
    const R *table;            // flame table (see DeviceFlame), copied to LDS
    const unsigned long long *masks; // [nxf] used-slot masks, copied to LDS
    void *buf;                 // cells * (1 + r) words
    DevStats *stats;
    unsigned long long *xf_count; // [nxf] selections by sorted position
    unsigned long long *bv_xfs;   // [bv_cap] sorted position of bad-value xforms
    double *bv_pts;               // [bv_cap][3]
    R lo[3], hi[3], mult_d[3];
    uint64_t mult_i[3];
    uint64_t cells;
    uint64_t seed;
    uint64_t walker_base;      // stream index of walker 0 of this launch
    uint64_t n_walkers;
    uint64_t batch;            // samples per walker
    uint64_t last_samples;     // samples of the last walker
    uint64_t bv_limit;
    uint64_t bv_cap;
    int nxf, n_slots, has_final, fin_nv, settle, r;
    int tab_reals;
    int xf0, fin0;
    // LDS cell cache (kernels built with it): cache_slots entries of
    // [key u32 | count u32 | r colour words], probed cache_ways deep; when
    // cache_direct the cell index is the slot (the whole buffer fits)
    int cache_slots, cache_log2, cache_ways, cache_direct;
    // slot and final-variation descriptors live in the kernel arguments so
    // the kernel reads them with scalar loads
    int slot_type[kMaxSlots];
    int slot_off[kMaxSlots];    // table offset of the slot's entries
    int slot_stride[kMaxSlots]; // entry size
    int fin_type[kMaxFinalVars];
    int fin_off[kMaxFinalVars];
    // xform record layout (flame_model.hpp XformLayout)
    int l_preA, l_preb, l_postA, l_postb, l_speed, l_color, l_has_color, l_has_pre,
        l_has_post, l_stride;
};

constexpr unsigned kEmptyKey = 0xFFFFFFFFu;

// byte offset of the LDS cell cache (after the flame table, the used-slot
// masks and the per-xform counters), 8-byte aligned
__host__ __device__ inline size_t cache_lds_offset(int tab_reals, int real_bytes, int nxf)
{
    size_t o = (size_t)tab_reals * (size_t)real_bytes + 8u * (size_t)nxf + 4u * (size_t)nxf;
    return (o + 7u) & ~(size_t)7u;
}

// bytes of one cell-cache entry: key, count and the colour words
__host__ __device__ inline size_t cache_slot_bytes(int r, int colour_bytes)
{
    return 8u + (size_t)r * (size_t)colour_bytes;
}

// precision modes (num_t / hist_t / colour word)
enum { MODE_F64_U64 = 0, MODE_F32_U32 = 1, MODE_F32_U64 = 2 };

} // namespace ffr

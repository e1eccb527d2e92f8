// C ABI of the renderer (include/ffr.h): flame parsing, device buffers,
// kernel launches and statistics.
#include "../../include/ffr.h"

#include "flame_model.hpp"
#include "rtc.hpp"
#include "util_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

using namespace ffr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(FFR_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Colour flames switch to binned accumulation after a render of at least
// kBinColourMinSamples samples made kBinColourAtomics or more global atomic
// adds (LDS-cache misses, a count and each colour word) per sample.  Measured
// on 64 batch flames of 2048^2 cells with three colours, 1000 samples per
// cell, FP32: binned renders faster from 0.55-0.6 atomics per sample on
// (1.03-1.42x at 0.6-0.9, 2-4x from 1.1 on), slower below 0.5 (0.3x for
// flames that stay in the cache).
constexpr uint64_t kBinColourMinSamples = 1ull << 26;
constexpr double kBinColourAtomics = 0.6;

// the kernels take the xform record offsets as constants (XformLay<D>);
// the table builder must lay records out the same way
template <int D> bool layout_matches(const XformLayout &l)
{
    using K = XformLay<D>;
    return l.preA == K::preA && l.preb == K::preb && l.postA == K::postA &&
           l.postb == K::postb && l.speed == K::speed && l.has_color == K::has_color &&
           l.has_pre == K::has_pre && l.has_post == K::has_post && l.color == K::color;
}
bool layout_matches_kernels(const XformLayout &l, int dims)
{
    return dims == 1 ? layout_matches<1>(l) : dims == 2 ? layout_matches<2>(l)
                                             : dims == 3 && layout_matches<3>(l);
}

struct DevTables {
    void *table = nullptr;
    unsigned long long *masks = nullptr;
    size_t table_bytes = 0;
    int tab_reals = 0;
};

} // namespace

struct ffr_renderer {
    ffr_options opt{};
    FlameSpec spec;
    int mode = MODE_F64_U64;
    int device = 0;
    int cu_count = 256;
    DeviceFlame<double> f64;
    DeviceFlame<float> f32;
    DevTables tab;
    int vs = 2;
    // the wave-per-walker kernel applies: linear variations only (at most
    // kChainSlots per xform), at most four colour dimensions
    bool chain_ok = false;
    bool chain_post_id = false; // ... and every post-affine is the identity
    bool chain_shared = false;  // ... and one pre-affine matrix and weight
    bool chain_diag = false;    // ... and that matrix is diagonal
    size_t chain_lds = 0;       // its LDS bytes
    int chain_per_cu = 0;       // FFR_CHAIN_PER_CU: its walkers-per-CU limit (0: the build's)
    size_t lds_bytes = 0;
    // buffer
    void *dbuf = nullptr;
    uint64_t cells = 0;
    uint32_t cell_words = 1;
    uint32_t word_bytes = 8;
    // statistics
    DevStats *dstats = nullptr;
    unsigned long long *dxf = nullptr;
    unsigned long long *dbv_xfs = nullptr;
    double *dbv_pts = nullptr;
    unsigned long long *dhist = nullptr;
    void *dargs = nullptr;                // the walker kernel's arguments
    unsigned *dhot = nullptr;             // hot-cell hints of the LDS cache
    // streamed buffer I/O: two pinned host chunks, two device chunks
    void *pin[2] = {nullptr, nullptr};
    void *dchunk[2] = {nullptr, nullptr};
    hipEvent_t io_ev[2] = {nullptr, nullptr};
    size_t io_chunk = 0;
    void *dimg = nullptr;                 // image pixels (allocated on first use)
    size_t dimg_bytes = 0;
    std::vector<unsigned char> hargs;     // their host copy, alive until the copy ran
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the binned walker's probe of the compiled kernel: a pass each
    hipEvent_t ev_ab[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t next_walker = 0;
    // cumulative statistics (getters of the reference)
    uint64_t s_iter = 0, s_plot = 0;
    std::vector<uint64_t> xf_dist; // by xform id
    double pmin[3], pmax[3];
    // last render
    uint64_t last_bad = 0, last_walkers = 0, last_batch = 0, last_atomics = 0,
             last_flush_atomics = 0;
    double last_ms = 0;
    std::vector<uint64_t> bv_ids;
    std::vector<double> bv_pts;
    bool pending = false;
    bool timed = false; // ev0/ev1 bracket the pending render (a batch times itself)
    int nxf = 0;
    std::vector<int> order; // sorted position -> xform id
    // launch policy: LDS cell cache (-1 auto, 0 off, 1 on), its LDS budget in
    // KiB, block size (0 auto); FFR_CACHE / FFR_CACHE_KB / FFR_BLOCK override
    int cache_pref = -1;
    int cache_kb = 144;
    int block_pref = 0;
    size_t max_lds = 65536;
    size_t lds_per_cu = 65536;
    int last_cache = 0;
    bool last_chain = false;
    // binned accumulation (-1 chosen per render, 0 off, 1 on where allowed);
    // FFR_BINNED overrides.  Its records, walker states and per-pass
    // arguments are allocated on the first binned render and kept.
    int bin_pref = -1;
    BinScratch bin;
    size_t bin_pay_bytes = 0; // bytes of bin.pay (and of bin.pay_sorted)
    // a colour flame whose renders made at least kBinColourAtomics global
    // atomic adds per sample: its later single renders take binned
    // accumulation (binned -1)
    bool bin_colour = false;
    size_t bin_hist_words = 0;
    size_t bin_tiles = 0;
    void *wstate = nullptr;
    size_t wstate_bytes = 0;
    void *dargs_bin = nullptr;
    size_t dargs_bin_n = 0;
    bool last_bin = false;
    // the flame compiled into its own kernel (-1 chosen by a probe, 0 off,
    // 1 on); FFR_COMPILED overrides.  rtc_fn: the kernel for the coming
    // launch (nullptr: the table-driven one); compiled_choice: the probe's
    // verdict per launch geometry (cache on/off, lanes per workgroup)
    int compiled_pref = -1;
    void *rtc_fn = nullptr;
    bool last_compiled = false;
    std::unordered_map<int, bool> compiled_choice;
    // cell_cache -1: the probe's verdict per walker count (true: cache kept)
    std::unordered_map<uint64_t, bool> cache_choice;
    // renders launched and samples asked for, over the renderer's life
    uint64_t launches = 0;
    uint64_t samples_asked = 0;
};

static int pick_variation_set(const uint32_t ts[4])
{
    bool linear_only = true, basic = true;
    for (int t = 0; t < kNumVariations; ++t) {
        if (!(ts[t / 32] >> (t % 32) & 1u))
            continue;
        if (t != 0)
            linear_only = false;
        if (!(t == 0 || t == 1 || t == 2 || t == 20 || t == 21 || t == 22))
            basic = false;
    }
    return linear_only ? 0 : (basic ? 1 : 2);
}

template <typename R>
static int setup_tables(ffr_renderer *r, const DeviceFlame<R> &d)
{
    if ((int)d.slot_type.size() > kMaxSlots)
        return fail(FFR_E_UNSUPPORTED, "too many variation slots");
    if ((int)d.fin_type.size() > kMaxFinalVars)
        return fail(FFR_E_UNSUPPORTED, "final xform has more than " +
                                           std::to_string(kMaxFinalVars) + " variations");
    if (d.nxf > kMaxXforms)
        return fail(FFR_E_UNSUPPORTED, "more than " + std::to_string(kMaxXforms) + " xforms");
    r->tab.tab_reals = (int)d.table.size();
    r->tab.table_bytes = d.table.size() * sizeof(R);
    HIP_TRY(hipMalloc(&r->tab.table, std::max<size_t>(r->tab.table_bytes, 16)));
    HIP_TRY(hipMemcpy(r->tab.table, d.table.data(), r->tab.table_bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&r->tab.masks, sizeof(unsigned long long) * (size_t)d.nxf));
    HIP_TRY(hipMemcpy(r->tab.masks, d.masks.data(), sizeof(unsigned long long) * (size_t)d.nxf,
                      hipMemcpyHostToDevice));
    r->lds_bytes = cache_lds_offset(r->tab.tab_reals, (int)sizeof(R), d.nxf);
    int max_lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  r->device));
    if (r->lds_bytes > (size_t)max_lds)
        return fail(FFR_E_UNSUPPORTED, "flame table of " + std::to_string(r->lds_bytes) +
                                           " bytes exceeds the " + std::to_string(max_lds) +
                                           "-byte LDS of a workgroup");
    r->vs = pick_variation_set(d.type_set);
    r->chain_ok = r->vs == 0 /* linear only, final xform included */ && d.r <= 4 &&
                  d.slot_type.size() <= (size_t)kChainSlots;
    // every xform's post-affine is exactly the identity (entries 1.0 and
    // +0.0, bit for bit): the wave-per-walker kernel then leaves it out (it
    // changes a finite point only in the signs of zeros)
    r->chain_post_id = true;
    for (int x = 0; x < d.nxf && r->chain_post_id; ++x) {
        const R *X = d.table.data() + d.xf0 + (size_t)x * (size_t)d.lay.stride;
        for (int i = 0; i < d.dims * d.dims; ++i) {
            const R v = X[d.lay.postA + i];
            const R want = (i / d.dims == i % d.dims) ? (R)1 : (R)0;
            if (!(v == want && !std::signbit(v)))
                r->chain_post_id = false;
        }
        for (int i = 0; i < d.dims; ++i)
            if (!(X[d.lay.postb + i] == (R)0 && !std::signbit(X[d.lay.postb + i])))
                r->chain_post_id = false;
    }
    // every xform applies the same pre-affine matrix (bit for bit, and in
    // 3-D every one applies it) and one variation of the same weight: the
    // kernel holds them once and broadcasts only the offsets
    r->chain_shared = d.slot_type.size() == 1 && d.nxf > 0;
    for (int x = 0; x < d.nxf && r->chain_shared; ++x) {
        const R *X = d.table.data() + d.xf0 + (size_t)x * (size_t)d.lay.stride;
        const R *X0 = d.table.data() + d.xf0;
        if (std::memcmp(X + d.lay.preA, X0 + d.lay.preA, sizeof(R) * (size_t)(d.dims * d.dims)))
            r->chain_shared = false;
        if (d.dims == 3 && X[d.lay.has_pre] == (R)0)
            r->chain_shared = false;
        if (!(d.masks[(size_t)x] & 1ull))
            r->chain_shared = false;
        const R *w = d.table.data() + d.slot_off[0] + (size_t)x * (size_t)d.slot_stride[0];
        if (std::memcmp(w, d.table.data() + d.slot_off[0], sizeof(R)))
            r->chain_shared = false;
    }
    // the shared matrix is diagonal (off-diagonal entries +0 or -0): the
    // kernel leaves out their products, which are zeros for a finite point
    r->chain_diag = r->chain_shared;
    for (int i = 0; i < d.dims * d.dims && r->chain_diag; ++i)
        if (i / d.dims != i % d.dims && d.table[(size_t)(d.xf0 + d.lay.preA + i)] != (R)0)
            r->chain_diag = false;
    // its LDS: the table, then kChainRows step rows per wave of the flame's
    // build
    const int slots = (int)d.slot_type.size();
    r->chain_lds =
        chain_stage_offset(r->tab.tab_reals, (int)sizeof(R), d.nxf) +
        (size_t)kChainWaves * (size_t)kChainRows * sizeof(R) *
            (size_t)chain_row_reals((int)sizeof(R), d.dims, d.r > 0 ? 4 : 0, chain_build_ns(slots),
                                    r->chain_post_id,
                                    chain_build_sh(r->chain_shared, r->chain_post_id, slots));
    if (r->chain_lds > (size_t)max_lds)
        r->chain_ok = false;
    r->nxf = d.nxf;
    r->order = d.order;
    return FFR_OK;
}

extern "C" {

void ffr_default_options(ffr_options *opt)
{
    if (!opt)
        return;
    std::memset(opt, 0, sizeof *opt);
    opt->precision = FFR_FP64;
    opt->word = FFR_WORD64;
    opt->settle_iters = -1;
    opt->bad_value_limit = 256;
    opt->bad_value_cap = 4096;
    opt->cell_cache = -1;
    opt->binned = -1;
    opt->compiled = -1;
}

const char *ffr_last_error(void) { return g_err.c_str(); }

int ffr_create(const char *flame_json, size_t len, const ffr_options *opt_in,
               ffr_renderer **out)
{
    if (!flame_json || !out)
        return fail(FFR_E_ARG, "ffr_create: null argument");
    *out = nullptr;
    ffr_options opt;
    if (opt_in)
        opt = *opt_in;
    else
        ffr_default_options(&opt);
    if (opt.precision > FFR_FP32 || opt.word > FFR_WORD32)
        return fail(FFR_E_ARG, "ffr_create: unknown precision or word option");
    if (opt.precision == FFR_FP64 && opt.word == FFR_WORD32)
        return fail(FFR_E_UNSUPPORTED,
                    "FP64 trajectories need 64-bit buffer words (sizeof(num_t) == sizeof(hist_t))");
    if (opt.bad_value_cap == 0)
        opt.bad_value_cap = 4096;
    std::unique_ptr<ffr_renderer> r(new ffr_renderer());
    r->opt = opt;
    r->device = opt.device;
    r->mode = opt.precision == FFR_FP64 ? MODE_F64_U64
                                        : (opt.word == FFR_WORD32 ? MODE_F32_U32 : MODE_F32_U64);
    const bool fp32 = opt.precision == FFR_FP32;
    try {
        Json doc = Json::parse(std::string(flame_json, len));
        r->spec = parse_flame(doc, fp32);
        if (fp32)
            r->f32 = compile_flame<float>(r->spec);
        else
            r->f64 = compile_flame<double>(r->spec);
    } catch (const std::exception &e) {
        return fail(FFR_E_FLAME, e.what());
    }
    {
        const XformLayout &lay = fp32 ? r->f32.lay : r->f64.lay;
        if (!layout_matches_kernels(lay, r->spec.dims))
            return fail(FFR_E_UNSUPPORTED, "ffr_create: xform record layout differs from the kernels'");
    }
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY(hipDeviceGetAttribute(&r->cu_count, hipDeviceAttributeMultiprocessorCount,
                                  r->device));
    {
        int v = 0;
        HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, r->device));
        r->max_lds = (size_t)v;
        v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  r->device) == hipSuccess && v > 0)
            r->lds_per_cu = (size_t)v;
        else
            r->lds_per_cu = r->max_lds;
        r->cache_pref = opt.cell_cache < 0 ? -1 : (opt.cell_cache ? 1 : 0);
        if (opt.cache_kib > 0)
            r->cache_kb = opt.cache_kib;
        if (opt.block > 0)
            r->block_pref = opt.block;
        if (const char *e = std::getenv("FFR_CACHE"))
            r->cache_pref = std::atoi(e);
        if (const char *e = std::getenv("FFR_CACHE_KB"))
            r->cache_kb = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("FFR_BLOCK"))
            r->block_pref = std::atoi(e);
        r->bin_pref = opt.binned < 0 ? -1 : (opt.binned ? 1 : 0);
        if (const char *e = std::getenv("FFR_BINNED"))
            r->bin_pref = std::atoi(e);
        r->compiled_pref = opt.compiled < 0 ? -1 : (opt.compiled ? 1 : 0);
        if (const char *e = std::getenv("FFR_CHAIN_PER_CU"))
            r->chain_per_cu = std::atoi(e);
        if (const char *e = std::getenv("FFR_COMPILED"))
            r->compiled_pref = std::atoi(e) < 0 ? -1 : (std::atoi(e) ? 1 : 0);
    }
    int rc = fp32 ? setup_tables(r.get(), r->f32) : setup_tables(r.get(), r->f64);
    if (rc != FFR_OK) {
        ffr_destroy(r.release());
        return rc;
    }
    r->cells = r->spec.cells();
    r->cell_words = 1u + (uint32_t)r->spec.color_dims;
    r->word_bytes = r->mode == MODE_F32_U32 ? 4u : 8u;
    const size_t bytes = (size_t)r->cells * r->cell_words * r->word_bytes;
    hipError_t e = hipMalloc(&r->dbuf, bytes);
    if (e != hipSuccess) {
        ffr_destroy(r.release());
        return fail(FFR_E_DEVICE, "hipMalloc of the " + std::to_string(bytes) +
                                      "-byte buffer: " + hipGetErrorString(e));
    }
    ffr_renderer *rp = r.release();
#define HIP_TRY_D(call)                                                                \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ffr_destroy(rp);                                                           \
            return fail(FFR_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                              \
    } while (0)
    HIP_TRY_D(hipStreamCreateWithFlags(&rp->own_stream, hipStreamNonBlocking));
    rp->stream = rp->own_stream;
    HIP_TRY_D(hipEventCreate(&rp->ev0));
    HIP_TRY_D(hipEventCreate(&rp->ev1));
    HIP_TRY_D(hipMalloc(&rp->dstats, sizeof(DevStats)));
    HIP_TRY_D(hipMalloc(&rp->dxf, sizeof(unsigned long long) * (size_t)std::max(rp->nxf, 1)));
    HIP_TRY_D(hipMalloc(&rp->dbv_xfs, sizeof(unsigned long long) * rp->opt.bad_value_cap));
    HIP_TRY_D(hipMalloc(&rp->dbv_pts, sizeof(double) * 3 * rp->opt.bad_value_cap));
    HIP_TRY_D(hipMalloc(&rp->dhist, sizeof(unsigned long long) * 3));
    HIP_TRY_D(hipMalloc(&rp->dargs, std::max(sizeof(WalkArgs<double>), sizeof(WalkArgs<float>))));
    // hints start empty: a tag no geometry has
    HIP_TRY_D(hipMalloc(&rp->dhot, sizeof(unsigned) * (1 + (size_t)kHotSlots)));
    HIP_TRY_D(hipMemsetAsync(rp->dhot, 0xFF, sizeof(unsigned) * (1 + (size_t)kHotSlots),
                             rp->stream));
    HIP_TRY_D((hipError_t)launch_fill_zero(rp->dbuf, bytes, rp->stream));
    HIP_TRY_D(hipStreamSynchronize(rp->stream));
#undef HIP_TRY_D
    rp->xf_dist.assign(rp->spec.xforms.size(), 0);
    for (int i = 0; i < 3; ++i) {
        rp->pmin[i] = INFINITY;
        rp->pmax[i] = -INFINITY;
    }
    *out = rp;
    return FFR_OK;
}

void ffr_destroy(ffr_renderer *r)
{
    if (!r)
        return;
    // teardown: errors here have nowhere to go
    (void)hipSetDevice(r->device);
    if (r->own_stream)
        (void)hipStreamSynchronize(r->own_stream);
    if (r->stream && r->stream != r->own_stream)
        (void)hipStreamSynchronize(r->stream);
    (void)hipFree(r->tab.table);
    (void)hipFree(r->tab.masks);
    (void)hipFree(r->dbuf);
    (void)hipFree(r->dstats);
    (void)hipFree(r->dxf);
    (void)hipFree(r->dbv_xfs);
    (void)hipFree(r->dbv_pts);
    (void)hipFree(r->dhist);
    (void)hipFree(r->dargs);
    (void)hipFree(r->dimg);
    (void)hipFree(r->dhot);
    (void)hipFree(r->bin.rec);
    (void)hipFree(r->bin.sorted);
    (void)hipFree(r->bin.hist);
    (void)hipFree(r->bin.total);
    (void)hipFree(r->bin.start);
    (void)hipFree(r->bin.pay);
    (void)hipFree(r->bin.pay_sorted);
    (void)hipFree(r->wstate);
    (void)hipFree(r->dargs_bin);
    for (int k = 0; k < 2; ++k) {
        if (r->pin[k])
            (void)hipHostFree(r->pin[k]);
        (void)hipFree(r->dchunk[k]);
        if (r->io_ev[k])
            (void)hipEventDestroy(r->io_ev[k]);
    }
    if (r->ev0)
        (void)hipEventDestroy(r->ev0);
    if (r->ev1)
        (void)hipEventDestroy(r->ev1);
    for (hipEvent_t ev : r->ev_ab)
        if (ev)
            (void)hipEventDestroy(ev);
    if (r->own_stream)
        (void)hipStreamDestroy(r->own_stream);
    delete r;
}

uint64_t ffr_auto_batch(const ffr_renderer *r, uint64_t samples)
{
    (void)r;
    // enough walkers to fill the device (about 2^19 lanes), each long
    // enough that its settle iterations stay a small share of its work
    const uint64_t target_walkers = 1ull << 19;
    uint64_t b = (samples + target_walkers - 1) / target_walkers;
    if (b < 1024)
        b = 1024;
    return b;
}

} // extern "C"

template <typename R>
static void fill_args(ffr_renderer *r, const DeviceFlame<R> &d, WalkArgs<R> &a)
{
    std::memset(&a, 0, sizeof a);
    a.table = (const R *)r->tab.table;
    a.masks = r->tab.masks;
    a.buf = r->dbuf;
    a.stats = r->dstats;
    a.xf_count = r->dxf;
    a.bv_xfs = r->dbv_xfs;
    a.bv_pts = r->dbv_pts;
    for (int i = 0; i < 3; ++i) {
        a.lo[i] = d.lo[i];
        a.hi[i] = d.hi[i];
        a.mult_d[i] = d.mult_d[i];
        a.mult_i[i] = d.mult_i[i];
    }
    a.cells = d.cells;
    a.seed = r->opt.seed;
    a.bv_limit = r->opt.bad_value_limit;
    a.bv_cap = r->opt.bad_value_cap;
    a.nxf = d.nxf;
    a.n_slots = (int)d.slot_type.size();
    a.has_final = d.has_final;
    a.fin_nv = (int)d.fin_type.size();
    a.settle = r->opt.settle_iters >= 0 ? (int)r->opt.settle_iters : (sizeof(R) == 4 ? 24 : 53);
    a.r = d.r;
    a.tab_reals = (int)d.table.size();
    a.xf0 = d.xf0;
    a.fin0 = d.fin0;
    a.l_preA = d.lay.preA;
    a.l_preb = d.lay.preb;
    a.l_postA = d.lay.postA;
    a.l_postb = d.lay.postb;
    a.l_speed = d.lay.speed;
    a.l_color = d.lay.color;
    a.l_has_color = d.lay.has_color;
    a.l_has_pre = d.lay.has_pre;
    a.l_has_post = d.lay.has_post;
    a.l_stride = d.lay.stride;
    for (size_t s = 0; s < d.slot_type.size(); ++s) {
        a.slot_type[s] = d.slot_type[s];
        a.slot_off[s] = d.slot_off[s];
        a.slot_stride[s] = d.slot_stride[s];
    }
    a.bt0 = d.bt0;
    a.bt_stride = d.bt_stride;
    a.bt_nvmax = d.bt_nvmax;
    a.bt_union = 0;
    if (d.bt0 >= 0)
        for (int x = 0; x < d.nxf; ++x)
            a.bt_union |= (int)d.table[(size_t)(d.bt0 + x * d.bt_stride + 1)];
    for (size_t j = 0; j < d.fin_type.size(); ++j) {
        a.fin_type[j] = d.fin_type[j];
        a.fin_off[j] = d.fin_off[j];
    }
}

// launch shape: what the launch may use; results do not depend on it
struct LaunchShape {
    int cache = -1;     // LDS cell cache: -1 chosen, 0 off, 1 on
    int cache_kb = 144; // LDS budget of the cache per workgroup, KiB
    int block = 0;      // lanes per workgroup, 0 chosen
    int task_rounds = 1; // walkers per lane in one workgroup task
    int bin = -1;       // binned accumulation (single renders): -1 chosen, 0 off, 1 on
};

static LaunchShape renderer_shape(const ffr_renderer *r)
{
    LaunchShape sh;
    sh.bin = r->bin_pref;
    sh.cache = r->cache_pref;
    sh.cache_kb = r->cache_kb;
    sh.block = r->block_pref;
    return sh;
}

// launch geometry of one render pass (see plan_launch)
struct LaunchPlan {
    bool chain = false; // one walker per wave (chain_kernel)
    bool bin = false;   // binned accumulation (walk_bin_kernel + bin_kernels)
    bool cache = false;
    int block = 256;
    int grid = 1;
    size_t lds = 0;
    int slots = 0, log2s = 0, ways = 1, direct = 0;
    uint64_t task_walkers = 256;
};

static size_t table_lds(const ffr_renderer *r)
{
    const int rb = r->mode == MODE_F64_U64 ? 8 : 4;
    return cache_lds_offset(r->tab.tab_reals, rb, r->nxf);
}

// Accumulation through the LDS cell cache needs cell indices below 2^32,
// at most 4 colour words, and per-task counts that fit 32 bits (a task of
// task_walkers walkers adds at most task_walkers * batch to one cell).  The
// cache takes the LDS left after the flame table, up to the budget; a
// buffer that fits entirely is indexed directly.  `base_lds` is the LDS the
// launch must hold besides the cache (the largest flame table of a batch).
static LaunchPlan plan_launch(const ffr_renderer *r, uint64_t walkers, uint64_t batch,
                              const LaunchShape &sh, size_t base_lds = 0)
{
    LaunchPlan p;
    const size_t base = std::max(table_lds(r), base_lds);
    const int per_cu = r->chain_per_cu > 0
                           ? r->chain_per_cu
                           : chain_walkers_per_cu(r->chain_post_id, r->spec.color_dims);
    const bool chain = sh.cache < 0 && sh.block <= 0 && base_lds == 0 && r->chain_ok &&
                       walkers <= (uint64_t)r->cu_count * (uint64_t)per_cu;
    // binned accumulation: single renders whose walkers all fit one pass's
    // records, of count-only buffers under 2^28 cells or colour buffers (up
    // to four colour dimensions) under 2^25; chosen for count-only buffers
    // from 2^22 cells on when the lane-per-walker kernel runs (the samples
    // spread beyond what an LDS cache holds), for colour flames whose
    // earlier renders spent many global atomics per sample (bin_colour),
    // or when asked for
    const uint64_t tw_bin = (uint64_t)kBinWalkBlock;
    const int ncol = r->spec.color_dims;
    const bool bin_ok = base_lds == 0 && ncol <= 4 &&
                        r->cells < (ncol > 0 ? (1ull << 25) : (1ull << 28)) &&
                        (walkers + tw_bin - 1) / tw_bin * tw_bin <= kBinPassRecords;
    const bool bin_auto = sh.bin < 0 && !chain && sh.cache < 0 && sh.block <= 0 &&
                          (ncol == 0 ? r->cells >= (1ull << 22) : r->bin_colour);
    if (bin_ok && (sh.bin == 1 || bin_auto)) {
        p.bin = true;
        p.block = kBinWalkBlock;
        p.task_walkers = tw_bin;
        p.grid = (int)std::min<uint64_t>((walkers + tw_bin - 1) / tw_bin, 1u << 30);
        p.lds = base;
        (void)batch;
        return p;
    }
    // few walkers of a linear flame: one walker per wave, its chain split
    // into parallel draws and plots around a short serial part (unless the
    // cache or the workgroup size is forced)
    if (chain) {
        p.chain = true;
        p.block = 64 * kChainWaves;
        p.task_walkers = (uint64_t)kChainWaves;
        p.grid = (int)((walkers + kChainWaves - 1) / kChainWaves);
        p.lds = r->chain_lds; // the table, then each wave's 64 step rows
        return p;
    }
    const int colour_bytes = r->word_bytes;
    const size_t slot_bytes = cache_slot_bytes(r->spec.color_dims, colour_bytes);
    bool want = sh.cache != 0 && r->cells < 0xFFFFFFFFull && r->spec.color_dims <= 4;
    // with fewer walkers than one wave per CU the render is bound by the
    // latency of one walker's chain, which the cache lengthens (unless
    // forced on); from there on the cache keeps hot cells off the device's
    // atomics (measured: 64000 walkers on a 256^2 flame, 0.12 G samples/s
    // without it, 16 with it)
    if (sh.cache < 0 && walkers < (uint64_t)r->cu_count * 64u)
        want = false;
    const int bmax = cache_block_max(r->mode == MODE_F64_U64, r->vs);
    p.block = (sh.block > 0 && sh.block <= bmax) ? sh.block : bmax;
    const uint64_t rounds = (uint64_t)std::max(1, sh.task_rounds);
    if (want && (uint64_t)p.block * batch >= (1ull << 31))
        want = false;
    if (want) {
        // (256 bytes stay free for the batched kernel's static LDS)
        size_t budget = std::min<size_t>(r->max_lds - 256, (size_t)sh.cache_kb << 10);
        size_t avail = budget > base ? budget - base : 0;
        int log2s = 0;
        while (((size_t)2 << log2s) * slot_bytes <= avail && log2s < 30)
            ++log2s;
        if (((size_t)1 << log2s) * slot_bytes > avail)
            want = false;
        else {
            p.log2s = log2s;
            p.slots = 1 << log2s;
            p.ways = 4;
            if ((uint64_t)p.slots >= r->cells) {
                while (p.log2s > 0 && ((uint64_t)1 << (p.log2s - 1)) >= r->cells)
                    --p.log2s;
                p.slots = 1 << p.log2s;
                p.direct = 1;
                p.ways = 1;
            }
        }
    }
    p.cache = want;
    if (p.cache) {
        p.lds = base + (size_t)p.slots * slot_bytes;
        uint64_t rr = std::min<uint64_t>(rounds, ((1ull << 31) - 1) / ((uint64_t)p.block * batch));
        p.task_walkers = (uint64_t)p.block * std::max<uint64_t>(rr, 1);
        int per_cu = std::max<int>(1, (int)(r->lds_per_cu / std::max<size_t>(p.lds, 1)));
        uint64_t tasks = (walkers + p.task_walkers - 1) / p.task_walkers;
        p.grid = (int)std::max<uint64_t>(
            1, std::min<uint64_t>(tasks, (uint64_t)r->cu_count * (uint64_t)per_cu));
    } else {
        p.block = walkers >= (uint64_t)r->cu_count * 256 ? 256 : 64;
        if (sh.block > 0 && sh.block <= 256)
            p.block = sh.block;
        p.task_walkers = (uint64_t)p.block * rounds;
        uint64_t tasks = (walkers + p.task_walkers - 1) / p.task_walkers;
        p.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(tasks, 1u << 30));
        p.lds = base;
    }
    return p;
}

template <typename R>
static void walk_args(ffr_renderer *r, const DeviceFlame<R> &d, const LaunchPlan &p,
                      uint64_t walker0, uint64_t n_walkers, uint64_t batch, uint64_t last,
                      WalkArgs<R> &a, bool hints = false)
{
    fill_args(r, d, a);
    a.walker_base = walker0;
    a.n_walkers = n_walkers;
    a.batch = batch;
    a.last_samples = last;
    a.task_walkers = p.task_walkers;
    a.cache_slots = p.slots;
    a.cache_log2 = p.log2s;
    a.cache_ways = p.ways;
    a.cache_direct = p.direct;
    // hot-cell hints for hashed caches (a task keeps a slot as a hint when
    // it counted at least an eighth of its samples per slot there, and at
    // least hot_thresh); the tag names the cache geometry of the hints.
    // Only for flames that few workgroups render at once: workgroups seeded
    // alike miss the same cells, and their global atomics then queue on the
    // same addresses (the headline flame on all 256 CUs ran 3x slower).
    a.hot = nullptr;
    a.hot_tag = 0;
    a.hot_thresh = 2;
    if (hints && p.cache && !p.direct && p.slots <= kHotSlots && r->dhot) {
        a.hot = r->dhot;
        a.hot_tag = 0x40000000u | ((unsigned)p.log2s << 8) | (unsigned)p.ways;
    }
}

static int launch_walk_any(int mode, int dims, const WalkLaunch &L)
{
    switch (mode) {
    case MODE_F64_U64:
        return dims == 1 ? launch_walk_f64u64_d1(L) : dims == 2 ? launch_walk_f64u64_d2(L)
                                                                 : launch_walk_f64u64_d3(L);
    case MODE_F32_U32:
        return dims == 1 ? launch_walk_f32u32_d1(L) : dims == 2 ? launch_walk_f32u32_d2(L)
                                                                 : launch_walk_f32u32_d3(L);
    default:
        return dims == 1 ? launch_walk_f32u64_d1(L) : dims == 2 ? launch_walk_f32u64_d2(L)
                                                                 : launch_walk_f32u64_d3(L);
    }
}

// the colour-register class the walker kernel is built for (launch_walk)
static int nc_class(int r, bool cache)
{
    if (cache && r <= 4)
        return r == 0 ? 0 : 4;
    return r == 0 ? 0 : (r <= 4 ? 4 : 128);
}

static int launch_walk_any(int mode, int dims, const WalkLaunch &L);

// the table-driven walker kernel for a plan, its arguments at r->dargs
static int launch_table_kernel(ffr_renderer *r, const LaunchPlan &p, int nslots)
{
    WalkLaunch L;
    L.nc = r->spec.color_dims;
    L.vs = r->vs;
    L.chain = p.chain;
    L.chain_slots = nslots;
    L.chain_post_id = r->chain_post_id;
    L.chain_shared = r->chain_shared;
    L.chain_diag = r->chain_diag;
    L.cache = p.cache;
    L.grid = p.grid;
    L.block = p.block;
    L.lds = p.lds;
    L.stream = r->stream;
    L.args = r->dargs;
    return launch_walk_any(r->mode, r->spec.dims, L);
}

// the flame's compiled kernel, its arguments (the WalkDyn part) at r->dargs
static int launch_compiled_kernel(ffr_renderer *r, const LaunchPlan &p, void *fn)
{
    void *dargs = r->dargs;
    void *params[] = {&dargs};
    return (int)hipModuleLaunchKernel((hipFunction_t)fn, (unsigned)p.grid, 1, 1,
                                      (unsigned)p.block, 1, 1, (unsigned)p.lds, r->stream,
                                      params, nullptr);
}

// With compiled -1 a renderer in steady use (its second render or later,
// once its renders have asked for this many samples in all) compiles the
// flame's kernel and probes whether it beats the table-driven one: the
// compile (about half a second) pays off only over repeated renders.
constexpr uint64_t kCompiledProbeSamples = 1ull << 33;

// Time kernels for the coming render: n launch plans, each with the flame's
// compiled kernel fns[k] or the table-driven one (nullptr), on the render's
// walkers with at most 1024 samples each, into a scratch buffer: the
// renderer's buffer and stream positions are untouched, and the statistics
// the runs leave are reset before the render.  ms[k]: the better of two
// runs (one, when after the first every other kernel took more than
// stop_ratio times kernel 0's time; 0: always two).
template <typename R>
static int time_kernels(ffr_renderer *r, const DeviceFlame<R> &d, const LaunchPlan *plans,
                        void *const *fns, int n, uint64_t walker0, uint64_t walkers,
                        uint64_t batch, float *ms, float stop_ratio = 0.f)
{
    const uint64_t pb = std::min<uint64_t>(batch, 1024);
    void *scratch = nullptr;
    HIP_TRY(hipMalloc(&scratch, ffr_buffer_bytes(r)));
    std::vector<std::vector<unsigned char>> h((size_t)n);
    for (int k = 0; k < n; ++k) {
        WalkArgs<R> a;
        walk_args(r, d, plans[k], walker0, walkers, pb, pb, a);
        a.buf = scratch;
        a.bv_limit = ~0ull; // the runs never stop at the bad-value limit
        h[(size_t)k].resize(sizeof a);
        std::memcpy(h[(size_t)k].data(), &a, sizeof a);
        ms[k] = 1e30f;
    }
    hipError_t e = hipSuccess;
    for (int round = 0; round < 2 && e == hipSuccess; ++round) {
        if (round == 1 && stop_ratio > 0.f) {
            bool all_slower = true;
            for (int k = 1; k < n; ++k)
                all_slower = all_slower && ms[k] > stop_ratio * ms[0];
            if (all_slower)
                break;
        }
        for (int k = 0; k < n && e == hipSuccess; ++k) {
            e = hipMemcpyAsync(r->dargs, h[(size_t)k].data(), h[(size_t)k].size(),
                               hipMemcpyHostToDevice, r->stream);
            if (e == hipSuccess)
                e = launch_reset_stats(r->dstats, r->dxf, r->nxf, r->stream)
                        ? hipErrorLaunchFailure
                        : hipSuccess;
            if (e == hipSuccess)
                e = hipEventRecord(r->ev0, r->stream);
            if (e == hipSuccess)
                e = (hipError_t)(fns[k] ? launch_compiled_kernel(r, plans[k], fns[k])
                                        : launch_table_kernel(r, plans[k],
                                                              (int)d.slot_type.size()));
            if (e == hipSuccess)
                e = hipEventRecord(r->ev1, r->stream);
            if (e == hipSuccess)
                e = hipEventSynchronize(r->ev1);
            float t = 0;
            if (e == hipSuccess)
                e = hipEventElapsedTime(&t, r->ev0, r->ev1);
            if (e == hipSuccess)
                ms[k] = std::min(ms[k], t);
        }
    }
    const hipError_t ef = hipFree(scratch);
    if (e == hipSuccess)
        e = ef;
    if (e != hipSuccess)
        return fail(FFR_E_DEVICE, std::string("kernel probe: ") + hipGetErrorString(e));
    return FFR_OK;
}

// With the cell cache left to the renderer (cell_cache -1), a renderer in
// steady use (its second render or later, once its renders have asked for
// this many samples in all) with at most kCacheProbeWalkersPerCu walkers
// per CU times the cached plan the renderer chose against the same launch
// without the cache, and keeps the faster per walker count.  (The cache
// keeps hot cells off the device's atomics but lengthens each walker's
// chain, which bounds renders of few walkers: the Sierpinski triangle at
// 64 walkers per CU rendered faster without it.  With many walkers the
// atomics bound instead: the headline flame's 2048 walkers per CU ran 40
// times slower without the cache, so there is nothing to probe.)
constexpr uint64_t kCacheProbeSamples = 1ull << 28;
constexpr uint64_t kCacheProbeWalkersPerCu = 256;

template <typename R>
static int choose_cache(ffr_renderer *r, const DeviceFlame<R> &d, LaunchPlan &p,
                        uint64_t walker0, uint64_t walkers, uint64_t batch, uint64_t samples)
{
    if (r->cache_pref >= 0 || !p.cache || p.chain || p.bin)
        return FFR_OK;
    LaunchShape sh = renderer_shape(r);
    sh.cache = 0;
    const auto known = r->cache_choice.find(walkers);
    if (known != r->cache_choice.end()) {
        if (!known->second)
            p = plan_launch(r, walkers, batch, sh);
        return FFR_OK;
    }
    // FFR_CACHE_PROBE_SAMPLES lowers the threshold (tests probe small renders)
    uint64_t threshold = kCacheProbeSamples;
    if (const char *env = std::getenv("FFR_CACHE_PROBE_SAMPLES"))
        threshold = std::strtoull(env, nullptr, 10);
    if (r->launches == 0 || r->samples_asked + samples < threshold ||
        walkers > kCacheProbeWalkersPerCu * (uint64_t)r->cu_count)
        return FFR_OK;
    const LaunchPlan plans[2] = {p, plan_launch(r, walkers, batch, sh)};
    void *const fns[2] = {nullptr, nullptr};
    float ms[2];
    // (one run each when the launch without the cache is far slower)
    const int rc = time_kernels(r, d, plans, fns, 2, walker0, walkers, batch, ms, 1.5f);
    if (rc != FFR_OK)
        return rc;
    // the cache stays unless the launch without it is 3 % faster
    const bool keep = !(ms[1] < 0.97f * ms[0]);
    r->cache_choice[walkers] = keep;
    if (!keep)
        p = plans[1];
    return FFR_OK;
}

// the flame's compiled kernel for this plan (lane-per-walker kernel only),
// compiled on first use; r->rtc_fn stays nullptr when the plan takes
// another kernel, compiled kernels are off, or the probe chose the table
template <typename R>
static int prepare_compiled(ffr_renderer *r, const DeviceFlame<R> &d, const LaunchPlan &p,
                            uint64_t walker0, uint64_t walkers, uint64_t batch, uint64_t last,
                            uint64_t samples)
{
    r->rtc_fn = nullptr;
    if (!r->compiled_pref || p.chain || p.bin)
        return FFR_OK;
    const int key = (p.cache ? 1 : 0) | (p.block << 1);
    const auto known = r->compiled_choice.find(key);
    const bool probe = r->compiled_pref < 0 && known == r->compiled_choice.end();
    if (r->compiled_pref < 0) {
        if (known != r->compiled_choice.end() && !known->second)
            return FFR_OK;
        // FFR_COMPILED_PROBE_SAMPLES lowers the threshold (tests probe small
        // renders)
        uint64_t threshold = kCompiledProbeSamples;
        if (const char *env = std::getenv("FFR_COMPILED_PROBE_SAMPLES"))
            threshold = std::strtoull(env, nullptr, 10);
        if (probe && (r->launches == 0 || r->samples_asked + samples < threshold))
            return FFR_OK;
    }
    WalkArgs<R> a;
    walk_args(r, d, p, walker0, walkers, batch, last, a);
    RtcKernel k;
    k.mode = r->mode;
    k.dims = r->spec.dims;
    k.nc = nc_class(r->spec.color_dims, p.cache);
    k.vs = r->vs;
    k.cache = p.cache;
    k.block_max = p.cache ? cache_block_max(r->mode == MODE_F64_U64, r->vs) : 256;
    std::string err;
    double compile_s = 0;
    void *fn = nullptr;
    if (rtc_flame_kernel(a, d.cw, k, r->device, &fn, &compile_s, err) != 0)
        return fail(FFR_E_DEVICE, err);
    if (probe) {
        const LaunchPlan plans[2] = {p, p};
        void *const fns[2] = {nullptr, fn};
        float ms[2];
        const int rc = time_kernels(r, d, plans, fns, 2, walker0, walkers, batch, ms);
        if (rc != FFR_OK)
            return rc;
        const bool faster = ms[1] < ms[0];
        r->compiled_choice[key] = faster;
        if (!faster)
            return FFR_OK;
    }
    r->rtc_fn = fn;
    return FFR_OK;
}

template <typename R>
static int launch_single(ffr_renderer *r, const DeviceFlame<R> &d, const LaunchPlan &p,
                         uint64_t walker0, uint64_t walkers, uint64_t batch, uint64_t last)
{
    WalkArgs<R> a;
    walk_args(r, d, p, walker0, walkers, batch, last, a);
    r->hargs.resize(sizeof a);
    std::memcpy(r->hargs.data(), &a, sizeof a);
    // a compiled flame's kernel reads only the launch part (WalkDyn, at the
    // start of WalkArgs); its flame fields are constants
    const size_t nbytes = r->rtc_fn ? sizeof(WalkDyn<R>) : sizeof a;
    hipError_t e = hipMemcpyAsync(r->dargs, r->hargs.data(), nbytes, hipMemcpyHostToDevice,
                                  r->stream);
    if (e != hipSuccess)
        return (int)e;
    if (r->rtc_fn)
        return launch_compiled_kernel(r, p, r->rtc_fn);
    return launch_table_kernel(r, p, (int)d.slot_type.size());
}

// walkers and samples per walker of a render of `samples`
static void split_samples(const ffr_renderer *r, uint64_t samples, uint64_t &batch,
                          uint64_t &walkers, uint64_t &last)
{
    batch = r->opt.batch_size ? r->opt.batch_size : ffr_auto_batch(r, samples);
    walkers = (samples + batch - 1) / batch;
    last = samples - (walkers - 1) * batch;
}

// grow a device allocation to at least `bytes` (contents not kept)
static hipError_t ensure_device(void **p, size_t &have, size_t bytes)
{
    if (have >= bytes)
        return hipSuccess;
    (void)hipFree(*p);
    *p = nullptr;
    have = 0;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess)
        have = bytes;
    return e;
}

// Binned accumulation of one render: passes of `span` iterations of every
// walker (the walkers' points and stream positions carry over in wstate),
// each pass writing at most kBinPassRecords records, sorted by tile and
// added into the buffer before the next pass overwrites them.
template <typename R>
static int launch_binned(ffr_renderer *r, const DeviceFlame<R> &d, const LaunchPlan &p,
                         uint64_t walker0, uint64_t walkers, uint64_t batch, uint64_t last,
                         uint64_t samples)
{
    const uint64_t tw = p.task_walkers;
    const uint64_t slots = (walkers + tw - 1) / tw * tw; // walker slots of the tasks
    // FFR_BIN_PASS_RECORDS makes passes smaller (tests run many passes)
    uint64_t pass_records = kBinPassRecords;
    if (const char *env = std::getenv("FFR_BIN_PASS_RECORDS")) {
        const unsigned long long v = std::strtoull(env, nullptr, 10);
        if (v >= 1 && v < pass_records)
            pass_records = v;
    }
    const uint64_t span = std::max<uint64_t>(1, std::min<uint64_t>(batch, pass_records / slots));
    const uint64_t npass = (batch + span - 1) / span;
    const uint64_t cap = slots * span;
    const int ncol = r->spec.color_dims; // colour components of each record
    const int tile_bits = ncol > 0 ? kBinTileBitsColour : kBinTileBits;
    const size_t ntiles = (size_t)((r->cells + (1ull << tile_bits) - 1) >> tile_bits);
    hipError_t e = hipSuccess;
    size_t have_rec = (size_t)r->bin.capacity * 4, have_sorted = have_rec;
    if (have_rec < cap * 4) {
        e = ensure_device((void **)&r->bin.rec, have_rec, (size_t)cap * 4);
        if (e == hipSuccess)
            e = ensure_device((void **)&r->bin.sorted, have_sorted, (size_t)cap * 4);
        r->bin.capacity = e == hipSuccess ? cap : 0;
    }
    // colour flames: the records' colours, component i at i * capacity
    if (e == hipSuccess && ncol > 0) {
        const size_t need = (size_t)ncol * (size_t)r->bin.capacity * sizeof(R);
        size_t have_sorted_pay = r->bin_pay_bytes;
        e = ensure_device(&r->bin.pay, r->bin_pay_bytes, need);
        if (e == hipSuccess)
            e = ensure_device(&r->bin.pay_sorted, have_sorted_pay, need);
        if (e != hipSuccess)
            r->bin_pay_bytes = 0;
    }
    size_t hist_bytes = r->bin_hist_words * 4;
    if (e == hipSuccess)
        e = ensure_device((void **)&r->bin.hist, hist_bytes,
                          bin_scratch_hist_words(cap, (int)ntiles) * 4);
    r->bin_hist_words = hist_bytes / 4;
    // per tile: records and (after the scatter) accumulation chunks, then
    // the chunks' first work items (2 (ntiles + 1) words); the regions
    size_t tile_bytes = r->bin_tiles * 4;
    if (e == hipSuccess && tile_bytes < 2 * (ntiles + 1) * 4) {
        size_t t2 = tile_bytes;
        e = ensure_device((void **)&r->bin.total, tile_bytes, 2 * (ntiles + 1) * 4);
        if (e == hipSuccess)
            e = ensure_device((void **)&r->bin.start, t2, 2 * (ntiles + 1) * 4);
    }
    r->bin_tiles = e == hipSuccess ? std::max(r->bin_tiles, 2 * (ntiles + 1)) : 0;
    if (e == hipSuccess && npass > 1)
        e = ensure_device(&r->wstate, r->wstate_bytes, (size_t)slots * sizeof(WalkerState<R>));
    size_t args_bytes = r->dargs_bin_n * sizeof(WalkArgs<R>);
    if (e == hipSuccess)
        e = ensure_device(&r->dargs_bin, args_bytes, (size_t)npass * sizeof(WalkArgs<R>));
    r->dargs_bin_n = args_bytes / sizeof(WalkArgs<R>);
    if (e != hipSuccess)
        return (int)e;
    // one argument block per pass, copied at once (the host copy stays
    // alive until the copy ran)
    r->hargs.resize((size_t)npass * sizeof(WalkArgs<R>));
    for (uint64_t k = 0; k < npass; ++k) {
        WalkArgs<R> a;
        walk_args(r, d, p, walker0, walkers, batch, last, a);
        a.rec = r->bin.rec;
        a.wstate = r->wstate;
        a.pay = r->bin.pay;
        a.pay_stride = r->bin.capacity;
        a.it0 = k * span;
        a.it1 = std::min<uint64_t>(batch, a.it0 + span);
        std::memcpy(r->hargs.data() + k * sizeof a, &a, sizeof a);
    }
    e = hipMemcpyAsync(r->dargs_bin, r->hargs.data(), r->hargs.size(), hipMemcpyHostToDevice,
                       r->stream);
    if (e != hipSuccess)
        return (int)e;
    // The flame's compiled walker (compiled 1; with -1 as for the lane
    // kernel: a renderer in steady use compiles it, and a render of three
    // passes or more times its first pass with the table-driven walker and
    // its second with the compiled one, then keeps the faster for the rest
    // and for later renders).  Both give the same records.
    void *fn = nullptr;
    bool ab = false;
    const int key = 0x100 | (p.block << 9); // (apart from the lane kernel's keys)
    if (r->compiled_pref != 0) {
        const auto known = r->compiled_choice.find(key);
        uint64_t threshold = kCompiledProbeSamples;
        if (const char *env = std::getenv("FFR_COMPILED_PROBE_SAMPLES"))
            threshold = std::strtoull(env, nullptr, 10);
        const bool want =
            r->compiled_pref > 0 || (known != r->compiled_choice.end() && known->second);
        ab = r->compiled_pref < 0 && known == r->compiled_choice.end() && npass >= 3 &&
             r->launches > 0 && r->samples_asked + samples >= threshold;
        if (want || ab) {
            WalkArgs<R> a;
            walk_args(r, d, p, walker0, walkers, batch, last, a);
            RtcKernel k;
            k.mode = r->mode;
            k.dims = r->spec.dims;
            k.nc = ncol;
            k.vs = r->vs;
            k.block_max = kBinWalkBlock;
            k.bin = true;
            std::string err;
            double compile_s = 0;
            if (rtc_flame_kernel(a, d.cw, k, r->device, &fn, &compile_s, err) != 0)
                return fail(FFR_E_DEVICE, err);
        }
        if (ab)
            for (hipEvent_t &ev : r->ev_ab)
                if (!ev)
                    HIP_TRY(hipEventCreate(&ev));
    }
    void *use = ab ? nullptr : fn;
    for (uint64_t k = 0; k < npass; ++k) {
        // (the walker kernel writes every record of the pass)
        const uint64_t n_rec = slots * (std::min<uint64_t>(batch, (k + 1) * span) - k * span);
        const unsigned char *args = (const unsigned char *)r->dargs_bin + k * sizeof(WalkArgs<R>);
        if (ab && k == 1)
            use = fn;
        if (ab && k < 2)
            HIP_TRY(hipEventRecord(r->ev_ab[2 * k], r->stream));
        int rc;
        if (use) {
            const void *pa = args;
            void *params[] = {&pa};
            rc = (int)hipModuleLaunchKernel((hipFunction_t)use, (unsigned)p.grid, 1, 1,
                                            (unsigned)p.block, 1, 1, (unsigned)p.lds, r->stream,
                                            params, nullptr);
        } else {
            WalkLaunch L;
            L.nc = ncol;
            L.vs = r->vs;
            L.bin = true;
            L.grid = p.grid;
            L.block = p.block;
            L.lds = p.lds;
            L.stream = r->stream;
            L.args = args;
            rc = launch_walk_any(r->mode, r->spec.dims, L);
        }
        if (rc)
            return rc;
        if (ab && k < 2)
            HIP_TRY(hipEventRecord(r->ev_ab[2 * k + 1], r->stream));
        if (ab && k == 1) {
            float t_table = 0, t_compiled = 0;
            HIP_TRY(hipEventSynchronize(r->ev_ab[3]));
            HIP_TRY(hipEventElapsedTime(&t_table, r->ev_ab[0], r->ev_ab[1]));
            HIP_TRY(hipEventElapsedTime(&t_compiled, r->ev_ab[2], r->ev_ab[3]));
            const bool faster = t_compiled < t_table;
            r->compiled_choice[key] = faster;
            use = faster ? fn : nullptr;
            ab = false;
        }
        rc = launch_bin_accumulate(r->bin, n_rec, r->cells, r->dbuf, (int)r->word_bytes, ncol,
                                   (int)sizeof(R), r->stream);
        if (rc)
            return rc;
    }
    r->last_compiled = use != nullptr;
    return 0;
}

static int launch(ffr_renderer *r, uint64_t samples)
{
    uint64_t batch, walkers, last;
    split_samples(r, samples, batch, walkers, last);
    LaunchPlan p = plan_launch(r, walkers, batch, renderer_shape(r));
    // (the cache and compiled-kernel probes run first: the reset after them
    // clears what they counted)
    int rc = r->mode == MODE_F64_U64
                 ? choose_cache(r, r->f64, p, r->next_walker, walkers, batch, samples)
                 : choose_cache(r, r->f32, p, r->next_walker, walkers, batch, samples);
    if (rc != FFR_OK)
        return rc;
    r->last_walkers = walkers;
    r->last_batch = batch;
    r->last_cache = p.cache ? p.slots : 0;
    r->last_chain = p.chain;
    r->last_bin = p.bin;
    rc = r->mode == MODE_F64_U64
             ? prepare_compiled(r, r->f64, p, r->next_walker, walkers, batch, last, samples)
             : prepare_compiled(r, r->f32, p, r->next_walker, walkers, batch, last, samples);
    if (rc != FFR_OK)
        return rc;
    r->last_compiled = r->rtc_fn != nullptr;
    rc = launch_reset_stats(r->dstats, r->dxf, r->nxf, r->stream);
    if (rc)
        return fail(FFR_E_DEVICE, std::string("stats reset: ") + hipGetErrorString((hipError_t)rc));
    if (p.bin) {
        static std::once_flag prepared;
        static int prep_rc = 0;
        std::call_once(prepared, [] { prep_rc = prepare_bin_kernels(); });
        if (prep_rc)
            return fail(FFR_E_DEVICE, std::string("binned accumulation: ") +
                                          hipGetErrorString((hipError_t)prep_rc));
    }
    HIP_TRY(hipEventRecord(r->ev0, r->stream));
    if (p.bin)
        rc = r->mode == MODE_F64_U64
                 ? launch_binned(r, r->f64, p, r->next_walker, walkers, batch, last, samples)
                 : launch_binned(r, r->f32, p, r->next_walker, walkers, batch, last, samples);
    else
        rc = r->mode == MODE_F64_U64
                 ? launch_single(r, r->f64, p, r->next_walker, walkers, batch, last)
                 : launch_single(r, r->f32, p, r->next_walker, walkers, batch, last);
    if (rc)
        return fail(FFR_E_DEVICE, std::string("walker kernel launch: ") +
                                      hipGetErrorString((hipError_t)rc));
    HIP_TRY(hipEventRecord(r->ev1, r->stream));
    r->next_walker += walkers;
    r->launches += 1;
    r->samples_asked += samples;
    r->pending = true;
    r->timed = true;
    return FFR_OK;
}

static int finish(ffr_renderer *r, ffr_stats *out)
{
    if (r->pending) {
        HIP_TRY(hipStreamSynchronize(r->stream));
        r->pending = false;
        DevStats st;
        HIP_TRY(hipMemcpy(&st, r->dstats, sizeof st, hipMemcpyDeviceToHost));
        std::vector<unsigned long long> xf((size_t)std::max(r->nxf, 1));
        HIP_TRY(hipMemcpy(xf.data(), r->dxf, sizeof(unsigned long long) * (size_t)r->nxf,
                          hipMemcpyDeviceToHost));
        if (r->timed) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, r->ev0, r->ev1));
            r->last_ms = ms;
        }
        r->s_iter += st.s_iter;
        r->s_plot += st.s_plot;
        for (int k = 0; k < r->nxf; ++k)
            r->xf_dist[(size_t)r->order[(size_t)k]] += xf[(size_t)k];
        for (int i = 0; i < r->spec.dims; ++i) {
            r->pmin[i] = std::min(r->pmin[i], st.pmin[i]);
            r->pmax[i] = std::max(r->pmax[i], st.pmax[i]);
        }
        r->last_bad = st.n_bad;
        r->last_atomics = st.atomics;
        r->last_flush_atomics = st.flush_atomics;
        // a colour flame whose samples missed the LDS cache often enough
        // renders faster binned from now on (its global atomics cost more
        // than the records' trips through HBM)
        if (r->spec.color_dims > 0 && !r->last_bin && !r->last_chain &&
            st.s_iter >= kBinColourMinSamples &&
            (double)(st.atomics - st.flush_atomics) >= kBinColourAtomics * (double)st.s_iter)
            r->bin_colour = true;
        size_t nrec = (size_t)std::min<uint64_t>(st.n_bad, r->opt.bad_value_cap);
        r->bv_ids.assign(nrec, 0);
        r->bv_pts.assign(nrec * 3, 0.0);
        if (nrec) {
            HIP_TRY(hipMemcpy(r->bv_ids.data(), r->dbv_xfs, sizeof(uint64_t) * nrec,
                              hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(r->bv_pts.data(), r->dbv_pts, sizeof(double) * 3 * nrec,
                              hipMemcpyDeviceToHost));
            for (auto &id : r->bv_ids)
                id = (uint64_t)r->order[(size_t)id];
        }
    }
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->s_iter = r->s_iter;
        out->s_plot = r->s_plot;
        out->n_bad = r->last_bad;
        out->walkers = r->last_walkers;
        out->batch_size = r->last_batch;
        for (int i = 0; i < 3; ++i) {
            out->pt_min[i] = i < r->spec.dims ? r->pmin[i] : 0.0;
            out->pt_max[i] = i < r->spec.dims ? r->pmax[i] : 0.0;
        }
        out->kernel_ms = r->last_ms;
        out->atomics = r->last_atomics;
        out->flush_atomics = r->last_flush_atomics;
        out->wave_per_walker = r->last_chain ? 1u : 0u;
        out->cache_slots = (uint32_t)r->last_cache;
        out->binned = r->last_bin ? 1u : 0u;
        out->compiled = r->last_compiled ? 1u : 0u;
    }
    if (r->last_bad > r->opt.bad_value_limit)
        return fail(FFR_E_BADVALUE_LIMIT, "render failure (flame may not be sufficiently contractive)");
    return FFR_OK;
}

extern "C" {

int ffr_render_async(ffr_renderer *r, uint64_t samples)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_render_async: null renderer");
    if (r->pending)
        return fail(FFR_E_ARG, "ffr_render_async: a render is already pending (call ffr_sync)");
    HIP_TRY(hipSetDevice(r->device));
    r->bv_ids.clear();
    r->bv_pts.clear();
    r->last_bad = 0;
    if (samples == 0)
        return FFR_OK;
    return launch(r, samples);
}

int ffr_sync(ffr_renderer *r, ffr_stats *stats)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_sync: null renderer");
    HIP_TRY(hipSetDevice(r->device));
    return finish(r, stats);
}

int ffr_render(ffr_renderer *r, uint64_t samples, ffr_stats *stats)
{
    int rc = ffr_render_async(r, samples);
    if (rc != FFR_OK)
        return rc;
    return ffr_sync(r, stats);
}

int ffr_set_stream(ffr_renderer *r, void *hip_stream)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_set_stream: null renderer");
    if (r->pending)
        return fail(FFR_E_ARG, "ffr_set_stream: a render is pending");
    r->stream = hip_stream ? (hipStream_t)hip_stream : r->own_stream;
    return FFR_OK;
}

void *ffr_device_buffer(ffr_renderer *r) { return r ? r->dbuf : nullptr; }
uint64_t ffr_buffer_cells(const ffr_renderer *r) { return r ? r->cells : 0; }
uint32_t ffr_cell_words(const ffr_renderer *r) { return r ? r->cell_words : 0; }
uint32_t ffr_word_bytes(const ffr_renderer *r) { return r ? r->word_bytes : 0; }
size_t ffr_buffer_bytes(const ffr_renderer *r)
{
    return r ? (size_t)r->cells * r->cell_words * r->word_bytes : 0;
}
int32_t ffr_dims(const ffr_renderer *r) { return r ? r->spec.dims : 0; }
int32_t ffr_color_dims(const ffr_renderer *r) { return r ? r->spec.color_dims : 0; }

int ffr_read_buffer(ffr_renderer *r, void *dst, size_t bytes)
{
    if (!r || !dst)
        return fail(FFR_E_ARG, "ffr_read_buffer: null argument");
    if (bytes != ffr_buffer_bytes(r))
        return fail(FFR_E_SIZE, "ffr_read_buffer: size mismatch");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY(hipStreamSynchronize(r->stream));
    HIP_TRY(hipMemcpy(dst, r->dbuf, bytes, hipMemcpyDeviceToHost));
    return FFR_OK;
}

int ffr_write_buffer(ffr_renderer *r, const void *src, size_t bytes)
{
    if (!r || !src)
        return fail(FFR_E_ARG, "ffr_write_buffer: null argument");
    if (bytes != ffr_buffer_bytes(r))
        return fail(FFR_E_SIZE, "ffr_write_buffer: size mismatch");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY(hipStreamSynchronize(r->stream));
    HIP_TRY(hipMemcpy(r->dbuf, src, bytes, hipMemcpyHostToDevice));
    return FFR_OK;
}

int ffr_add_buffer(ffr_renderer *r, const void *src, size_t bytes)
{
    if (!r || !src)
        return fail(FFR_E_ARG, "ffr_add_buffer: null argument");
    if (bytes != ffr_buffer_bytes(r))
        return fail(FFR_E_SIZE, "BufferRenderer::addBuffer(): sizes do not match");
    HIP_TRY(hipSetDevice(r->device));
    // staged in chunks of whole cells through one device scratch buffer
    const size_t cell_bytes = (size_t)r->cell_words * r->word_bytes;
    const size_t chunk_cells = std::max<size_t>(1, (size_t)(256u << 20) / cell_bytes);
    const size_t chunk = std::min(bytes, chunk_cells * cell_bytes);
    void *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, chunk));
    int rc = FFR_OK;
    for (size_t off = 0; off < bytes && rc == FFR_OK; off += chunk) {
        size_t n = std::min(chunk, bytes - off);
        hipError_t e = hipMemcpyAsync(tmp, (const char *)src + off, n, hipMemcpyHostToDevice,
                                      r->stream);
        if (e == hipSuccess)
            e = (hipError_t)launch_add_cells((char *)r->dbuf + off, tmp,
                                             n / r->word_bytes, (int)r->cell_words,
                                             (int)r->word_bytes, r->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess)
            rc = fail(FFR_E_DEVICE, std::string("ffr_add_buffer: ") + hipGetErrorString(e));
    }
    (void)hipFree(tmp);
    return rc;
}

int ffr_add_buffer_noalloc(ffr_renderer *r, const void *src, size_t bytes)
{
    if (!r || !src)
        return fail(FFR_E_ARG, "ffr_add_buffer_noalloc: null argument");
    if (bytes != ffr_buffer_bytes(r))
        return fail(FFR_E_SIZE, "BufferRenderer::addBuffer(): sizes do not match");
    const unsigned long long cell_bytes = (unsigned long long)r->cell_words * r->word_bytes;
    if (cell_bytes > kNoAllocScratchBytes)
        return fail(FFR_E_SIZE, "ffr_add_buffer_noalloc: a cell of " +
                                    std::to_string(cell_bytes) +
                                    " bytes exceeds the staging memory of " +
                                    std::to_string(kNoAllocScratchBytes) + " bytes");
    HIP_TRY(hipSetDevice(r->device));
    const int rc = add_cells_staged_noalloc(r->dbuf, src, bytes, (int)r->cell_words,
                                            (int)r->word_bytes, r->stream);
    if (rc)
        return fail(FFR_E_DEVICE, std::string("ffr_add_buffer_noalloc: ") +
                                      hipGetErrorString((hipError_t)rc));
    return FFR_OK;
}

} // extern "C"

namespace {

// the streamed I/O chunks: 32 MiB of whole cells (or the whole buffer),
// allocated on first use and kept until ffr_destroy
int io_init(ffr_renderer *r)
{
    if (r->io_chunk)
        return FFR_OK;
    const size_t cell_bytes = (size_t)r->cell_words * r->word_bytes;
    const size_t chunk =
        std::min(ffr_buffer_bytes(r), std::max<size_t>(1, (size_t)(32u << 20) / cell_bytes) *
                                          cell_bytes);
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipHostMalloc(&r->pin[k], chunk, hipHostMallocDefault));
        HIP_TRY(hipMalloc(&r->dchunk[k], chunk));
        HIP_TRY(hipEventCreateWithFlags(&r->io_ev[k], hipEventDisableTiming));
    }
    r->io_chunk = chunk;
    return FFR_OK;
}

// fill dst from the callback until n bytes or the end
size_t read_full(ffr_read_fn read, void *ctx, void *dst, size_t n)
{
    size_t got = 0;
    while (got < n) {
        const size_t k = read(ctx, (char *)dst + got, n - got);
        if (k == 0)
            break;
        got += std::min(k, n - got);
    }
    return got;
}

} // namespace

extern "C" {

int ffr_add_buffer_stream(ffr_renderer *r, ffr_read_fn read, void *ctx)
{
    if (!r || !read)
        return fail(FFR_E_ARG, "ffr_add_buffer_stream: null argument");
    if (r->pending)
        return fail(FFR_E_ARG, "ffr_add_buffer_stream: a render is pending (call ffr_sync)");
    HIP_TRY(hipSetDevice(r->device));
    const int rc0 = io_init(r);
    if (rc0 != FFR_OK)
        return rc0;
    const size_t bytes = ffr_buffer_bytes(r), chunk = r->io_chunk;
    // chunk i goes through pinned buffer i % 2: before it is refilled, the
    // add of chunk i - 2 (which read it through its device chunk) is waited for
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += chunk, ++i) {
        const int k = (int)(i & 1);
        const size_t n = std::min(chunk, bytes - off);
        if (i >= 2)
            HIP_TRY(hipEventSynchronize(r->io_ev[k]));
        if (read_full(read, ctx, r->pin[k], n) != n) {
            HIP_TRY(hipStreamSynchronize(r->stream));
            return fail(FFR_E_IO, "error reading file");
        }
        HIP_TRY(hipMemcpyAsync(r->dchunk[k], r->pin[k], n, hipMemcpyHostToDevice, r->stream));
        HIP_TRY((hipError_t)launch_add_cells((char *)r->dbuf + off, r->dchunk[k],
                                             n / r->word_bytes, (int)r->cell_words,
                                             (int)r->word_bytes, r->stream));
        HIP_TRY(hipEventRecord(r->io_ev[k], r->stream));
    }
    HIP_TRY(hipStreamSynchronize(r->stream));
    return FFR_OK;
}

int ffr_read_buffer_stream(ffr_renderer *r, ffr_write_fn write, void *ctx)
{
    if (!r || !write)
        return fail(FFR_E_ARG, "ffr_read_buffer_stream: null argument");
    if (r->pending)
        return fail(FFR_E_ARG, "ffr_read_buffer_stream: a render is pending (call ffr_sync)");
    HIP_TRY(hipSetDevice(r->device));
    const int rc0 = io_init(r);
    if (rc0 != FFR_OK)
        return rc0;
    const size_t bytes = ffr_buffer_bytes(r), chunk = r->io_chunk;
    const size_t nchunks = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t i) {
        const int k = (int)(i & 1);
        const size_t off = i * chunk, n = std::min(chunk, bytes - off);
        hipError_t e = hipMemcpyAsync(r->pin[k], (const char *)r->dbuf + off, n,
                                      hipMemcpyDeviceToHost, r->stream);
        if (e == hipSuccess)
            e = hipEventRecord(r->io_ev[k], r->stream);
        return e;
    };
    // the copy of chunk i + 1 runs while chunk i is written out; pinned
    // buffer (i + 1) % 2 held chunk i - 1, which is written out already
    HIP_TRY(issue(0));
    for (size_t i = 0; i < nchunks; ++i) {
        const int k = (int)(i & 1);
        if (i + 1 < nchunks)
            HIP_TRY(issue(i + 1));
        HIP_TRY(hipEventSynchronize(r->io_ev[k]));
        const size_t n = std::min(chunk, bytes - i * chunk);
        if (write(ctx, r->pin[k], n) != n) {
            HIP_TRY(hipStreamSynchronize(r->stream));
            return fail(FFR_E_IO, "error writing output");
        }
    }
    return FFR_OK;
}

int ffr_check_flame(const char *flame_json, size_t len, uint32_t precision, int32_t *ids,
                    double *cw, size_t cap, size_t *n_xforms)
{
    if (!flame_json)
        return fail(FFR_E_ARG, "ffr_check_flame: null flame");
    try {
        Json doc = Json::parse(std::string(flame_json, len));
        FlameSpec spec = parse_flame(doc, precision == FFR_FP32);
        std::vector<int> order;
        std::vector<double> cum;
        if (precision == FFR_FP32) {
            auto d = compile_flame<float>(spec);
            order = d.order;
            for (float v : d.cw)
                cum.push_back((double)v);
        } else {
            auto d = compile_flame<double>(spec);
            order = d.order;
            cum = d.cw;
        }
        for (size_t k = 0; k < order.size() && k < cap; ++k) {
            if (ids)
                ids[k] = order[k];
            if (cw)
                cw[k] = cum[k];
        }
        if (n_xforms)
            *n_xforms = order.size();
    } catch (const std::exception &e) {
        return fail(FFR_E_FLAME, e.what());
    }
    return FFR_OK;
}

int ffr_histogram_l1(ffr_renderer *a, ffr_renderer *b, double *l1)
{
    if (!a || !b || !l1)
        return fail(FFR_E_ARG, "ffr_histogram_l1: null argument");
    if (a->cells != b->cells || a->device != b->device)
        return fail(FFR_E_SIZE, "ffr_histogram_l1: buffers of different shapes or devices");
    uint64_t sa = 0, sb = 0, mn, mx;
    int rc = ffr_histogram_stats(a, &sa, &mn, &mx);
    if (rc)
        return rc;
    rc = ffr_histogram_stats(b, &sb, &mn, &mx);
    if (rc)
        return rc;
    HIP_TRY(hipSetDevice(a->device));
    double *d = reinterpret_cast<double *>(a->dhist); // 3 words of scratch
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY((hipError_t)launch_l1_counts(a->dbuf, (int)a->cell_words, (int)a->word_bytes, b->dbuf,
                                         (int)b->cell_words, (int)b->word_bytes, a->cells,
                                         sa ? 1.0 / (double)sa : 0.0, sb ? 1.0 / (double)sb : 0.0,
                                         d, a->stream));
    HIP_TRY(hipMemcpyAsync(l1, d, sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    return FFR_OK;
}

int ffr_choose_precision(const char *flame_json, size_t len, const ffr_options *opt,
                         uint64_t probe_samples, double threshold, uint32_t *precision,
                         double *l1)
{
    if (!flame_json || !precision)
        return fail(FFR_E_ARG, "ffr_choose_precision: null argument");
    ffr_options o;
    if (opt)
        o = *opt;
    else
        ffr_default_options(&o);
    o.word = FFR_WORD64;
    // both renders draw the same words at the same iteration: one settle
    // count for both (the defaults differ between the precisions)
    if (o.settle_iters < 0)
        o.settle_iters = 53;
    ffr_renderer *r32 = nullptr, *r64 = nullptr;
    o.precision = FFR_FP32;
    int rc = ffr_create(flame_json, len, &o, &r32);
    if (rc == FFR_OK) {
        o.precision = FFR_FP64;
        rc = ffr_create(flame_json, len, &o, &r64);
    }
    int rc32 = FFR_OK, rc64 = FFR_OK;
    if (rc == FFR_OK) {
        rc32 = ffr_render(r32, probe_samples, nullptr);
        rc64 = ffr_render(r64, probe_samples, nullptr);
        if (rc32 != FFR_OK && rc32 != FFR_E_BADVALUE_LIMIT)
            rc = rc32;
        else if (rc64 != FFR_OK && rc64 != FFR_E_BADVALUE_LIMIT)
            rc = rc64;
    }
    double d = 0.0;
    if (rc == FFR_OK)
        rc = ffr_histogram_l1(r32, r64, &d);
    ffr_destroy(r32);
    ffr_destroy(r64);
    if (rc != FFR_OK)
        return rc;
    // an FP32 render stopped by bad values where FP64 was not: the FP32
    // attractor escapes
    const bool escapes = rc32 == FFR_E_BADVALUE_LIMIT && rc64 == FFR_OK;
    *precision = (!escapes && d <= threshold) ? FFR_FP32 : FFR_FP64;
    if (l1)
        *l1 = d;
    return FFR_OK;
}

int ffr_clear_buffer(ffr_renderer *r)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_clear_buffer: null renderer");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY((hipError_t)launch_fill_zero(r->dbuf, ffr_buffer_bytes(r), r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return FFR_OK;
}

int ffr_histogram_stats(ffr_renderer *r, uint64_t *sum, uint64_t *mn, uint64_t *mx)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_histogram_stats: null renderer");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY((hipError_t)launch_hist_stats(r->dbuf, r->cells, (int)r->cell_words,
                                          (int)r->word_bytes, r->dhist, r->stream));
    unsigned long long h[3];
    HIP_TRY(hipMemcpyAsync(h, r->dhist, sizeof h, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    if (sum)
        *sum = h[0];
    if (mn)
        *mn = (r->word_bytes == 4 && h[1] == ~0ull) ? 0xffffffffull : h[1];
    if (mx)
        *mx = h[2];
    return FFR_OK;
}

int ffr_atomic_rate(int32_t device, uint64_t cells, uint32_t cell_words, uint32_t word_bytes,
                    uint64_t updates, double *atomics_per_s)
{
    if (!atomics_per_s || cells == 0 || cells >= (1ull << 32) || cell_words == 0 ||
        cell_words > 5 || (word_bytes != 4 && word_bytes != 8) || updates == 0)
        return fail(FFR_E_ARG, "ffr_atomic_rate: cells in [1, 2^32), 1-5 words of 4 or 8 bytes, "
                               "updates > 0");
    HIP_TRY(hipSetDevice(device));
    const size_t bytes = (size_t)cells * cell_words * word_bytes;
    void *buf = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FFR_OK;
    float ms = 0.f;
    auto step = [&](hipError_t e, const char *what) {
        if (rc == FFR_OK && e != hipSuccess)
            rc = fail(FFR_E_DEVICE, std::string("ffr_atomic_rate: ") + what + ": " +
                                        hipGetErrorString(e));
        return rc == FFR_OK;
    };
    if (step(hipMalloc(&buf, bytes), "hipMalloc") &&
        step(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") &&
        step(hipEventCreate(&e0), "event") && step(hipEventCreate(&e1), "event") &&
        step((hipError_t)launch_fill_zero(buf, bytes, s), "clear") &&
        step((hipError_t)launch_scatter_atomics(buf, cells, (int)cell_words, (int)word_bytes,
                                                std::min<uint64_t>(updates, 1u << 24), s),
             "warm-up") &&
        step(hipEventRecord(e0, s), "record") &&
        step((hipError_t)launch_scatter_atomics(buf, cells, (int)cell_words, (int)word_bytes,
                                                updates, s),
             "scatter") &&
        step(hipEventRecord(e1, s), "record") && step(hipStreamSynchronize(s), "sync") &&
        step(hipEventElapsedTime(&ms, e0, e1), "elapsed")) {
        *atomics_per_s = ms > 0.f ? (double)updates * cell_words / ((double)ms * 1e-3) : 0.0;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (s)
        (void)hipStreamDestroy(s);
    (void)hipFree(buf);
    return rc;
}

size_t ffr_image_bytes(const ffr_renderer *r, int mode, int bits)
{
    if (!r || r->spec.dims != 2 || mode < FFR_IMG_MONO || mode > FFR_IMG_COLOR ||
        (bits != 8 && bits != 16))
        return 0;
    const size_t per = mode == FFR_IMG_MONO ? 1 : (size_t)(bits / 8);
    return (size_t)r->cells * (mode == FFR_IMG_COLOR ? 3 : 1) * per;
}

namespace {

// both image entry points: dst is host memory (to_host) or a device address
int render_image(ffr_renderer *r, int mode, double gamma, int bits, void *dst, size_t bytes,
                 ffr_image_info *info, bool to_host)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_render_image: null renderer");
    if (r->pending)
        return fail(FFR_E_ARG, "ffr_render_image: a render is pending (call ffr_sync)");
    // the checks of ffr_img.cpp, in its order (:92-97, :130-134, :149-157,
    // :236-237, :290-291); gamma and the log bound compare against eps of num_t
    const bool f32 = r->word_bytes == 4;
    if (f32 ? (float)gamma < 1e-10f : gamma < 1e-20)
        return fail(FFR_E_ARG, "gamma too small");
    if (bits != 8 && bits != 16)
        return fail(FFR_E_ARG, "bits per channel must be 8 or 16");
    if (r->spec.dims != 2)
        return fail(FFR_E_UNSUPPORTED, "only 2D flames supported");
    if (mode < FFR_IMG_MONO || mode > FFR_IMG_COLOR)
        return fail(FFR_E_ARG,
                    "must choose exactly 1 coloring flag (--monochrome, --grayscale, --color)");
    const size_t need = ffr_image_bytes(r, mode, bits);
    if (dst && (to_host ? bytes != need : bytes < need))
        return fail(FFR_E_SIZE, "ffr_render_image: expected " + std::to_string(need) +
                                    " bytes of pixels");
    HIP_TRY(hipSetDevice(r->device));
    HIP_TRY((hipError_t)launch_hist_stats(r->dbuf, r->cells, (int)r->cell_words,
                                          (int)r->word_bytes, r->dhist, r->stream));
    unsigned long long h[3];
    HIP_TRY(hipMemcpyAsync(h, r->dhist, sizeof h, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    const uint64_t mn = (f32 && h[1] == ~0ull) ? 0xffffffffull : h[1], mx = h[2];
    // the log scaler's bounds (ffr_img.cpp:223-236), in num_t
    const double smin = f32 ? (double)std::log(1.0f + (float)mn) : std::log(1.0 + (double)mn);
    const double smax = f32 ? (double)std::log(1.0f + (float)mx) : std::log(1.0 + (double)mx);
    if (info) {
        info->hist_min = mn;
        info->hist_max = mx;
        info->scale_min = smin;
        info->scale_max = smax;
        info->width = (uint32_t)r->spec.size[0];
        info->height = (uint32_t)r->spec.size[1];
        info->channels = mode == FFR_IMG_COLOR ? 3u : 1u;
        info->bytes_per_channel = mode == FFR_IMG_MONO ? 1u : (uint32_t)(bits / 8);
        info->kernel_ms = 0.0;
    }
    if (f32 ? (float)smax < 1e-10f : smax < 1e-20)
        return fail(FFR_E_EMPTY, "histogram is (probably) empty");
    if (mode == FFR_IMG_COLOR && r->spec.color_dims != 3)
        return fail(FFR_E_UNSUPPORTED, "buffer must use 3 color dimensions");
    if (!dst)
        return FFR_OK;
    void *out = dst;
    if (to_host) {
        if (r->dimg_bytes < need) {
            HIP_TRY(hipFree(r->dimg));
            r->dimg = nullptr;
            r->dimg_bytes = 0;
            HIP_TRY(hipMalloc(&r->dimg, need));
            r->dimg_bytes = need;
        }
        out = r->dimg;
    }
    // num_t gp = 1.0 / arg_gamma (ffr_img.cpp:240): a double division, then num_t
    const double gp = f32 ? (double)(float)(1.0 / (double)(float)gamma) : 1.0 / gamma;
    HIP_TRY(hipEventRecord(r->ev0, r->stream));
    HIP_TRY((hipError_t)launch_tonemap(r->dbuf, r->cells, (int)r->cell_words,
                                       (int)r->word_bytes, mode, mx, gp,
                                       mode == FFR_IMG_MONO ? 8 : bits, out, r->stream));
    HIP_TRY(hipEventRecord(r->ev1, r->stream));
    if (to_host)
        HIP_TRY(hipMemcpyAsync(dst, r->dimg, need, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    if (info)
        info->kernel_ms = (double)ms;
    return FFR_OK;
}

} // namespace

int ffr_render_image(ffr_renderer *r, int mode, double gamma, int bits, void *pixels,
                     size_t bytes, ffr_image_info *info)
{
    return render_image(r, mode, gamma, bits, pixels, bytes, info, true);
}

int ffr_render_image_device(ffr_renderer *r, int mode, double gamma, int bits,
                            void *device_pixels, size_t bytes, ffr_image_info *info)
{
    return render_image(r, mode, gamma, bits, device_pixels, bytes, info, false);
}

size_t ffr_xform_id_count(const ffr_renderer *r) { return r ? r->xf_dist.size() : 0; }

int ffr_xform_distribution(const ffr_renderer *r, uint64_t *out, size_t n)
{
    if (!r || (!out && n))
        return fail(FFR_E_ARG, "ffr_xform_distribution: null argument");
    if (n != r->xf_dist.size())
        return fail(FFR_E_SIZE, "ffr_xform_distribution: expected " +
                                    std::to_string(r->xf_dist.size()) + " entries");
    std::copy(r->xf_dist.begin(), r->xf_dist.end(), out);
    return FFR_OK;
}

int ffr_bad_values(const ffr_renderer *r, uint64_t *ids, double *pts, size_t cap, size_t *n)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_bad_values: null renderer");
    size_t m = std::min(cap, r->bv_ids.size());
    const int d = r->spec.dims;
    for (size_t i = 0; i < m; ++i) {
        if (ids)
            ids[i] = r->bv_ids[i];
        if (pts)
            for (int k = 0; k < d; ++k)
                pts[i * (size_t)d + (size_t)k] = r->bv_pts[i * 3 + (size_t)k];
    }
    if (n)
        *n = r->bv_ids.size();
    return FFR_OK;
}

int ffr_xform_order(const ffr_renderer *r, int32_t *ids, double *cw, size_t n)
{
    if (!r)
        return fail(FFR_E_ARG, "ffr_xform_order: null renderer");
    if (n != (size_t)r->nxf)
        return fail(FFR_E_SIZE, "ffr_xform_order: expected " + std::to_string(r->nxf) +
                                    " entries");
    for (int k = 0; k < r->nxf; ++k) {
        if (ids)
            ids[k] = r->order[(size_t)k];
        if (cw)
            cw[k] = r->mode == MODE_F64_U64 ? r->f64.cw[(size_t)k] : (double)r->f32.cw[(size_t)k];
    }
    return FFR_OK;
}

} // extern "C"

#include "ffr_batch.inc"
#include "ffr_server.inc"

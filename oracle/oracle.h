/*
 * CPU oracle for the ffr-buf buffer render -- TEST INFRASTRUCTURE ONLY.
 *
 * This is an independent plain-C restatement of the reference renderer's hot
 * path (tkoz0/flame-fractal-renderer), used only by tests/, the smoke check and
 * the CPU-baseline leg of bench.py, to check results against.  Nothing in the product
 * links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   - BufferRenderer::_init / _render_batch / render / renderSeeded
 *       src/renderers/buffer_renderer.hpp:114-140, 150-250, 269-338, 349-372
 *   - RenderIterator::_init / iterate / _bad_value / _in_bounds
 *       src/renderers/render_iterator.hpp:52-79, 106-139
 *   - Flame::_optimize / _setupCumulativeWeights / getRandomXForm
 *       src/types/flame.hpp:44-83, 212-219
 *   - XForm::_optimize / applyIteration   src/types/xform.hpp:47-58, 211-227
 *   - Affine::apply_to                    src/types/affine.hpp:104-110
 *   - Point norms / angle                 src/types/point.hpp:228-356, 415-420
 *   - all 98 variations                   src/variations/variations.hpp:160-2376
 *   - FlameRNG conversions                src/rng/flame_rng.hpp:60-219
 *   - Isaac<word_t,4>                     src/rng/isaac.hpp:77-329
 *
 * The one deliberate change is the random stream: each walker (one reference
 * "batch") draws its words from its own Philox4x32-10 counter stream keyed by
 * (seed, walker index) instead of a clock-seeded thread-local ISAAC, so that a
 * render is reproducible and the GPU kernels can be compared to it bit for bit.
 * The ISAAC restatement is kept as a stand-alone generator and is checked
 * against the reference's own isaac.hpp.
 *
 * Parity status of each part is listed in the header of oracle.c.
 */
#ifndef FFR_ORACLE_H
#define FFR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_PARAMS 16
#define ORC_MAX_DIMS 3

/* one variation as written in the flame JSON (raw doubles, catalogue order) */
typedef struct {
    int32_t type;      /* catalogue id, see oracle/model.py VARIATIONS */
    int32_t axis_x;    /* 2-D variations in 3-D flames: axes (else -1) */
    int32_t axis_y;
    int32_t n_params;
    double weight;
    double params[ORC_MAX_PARAMS];
} orc_var;

/* one xform as written in the flame JSON */
typedef struct {
    double weight;     /* unused for the final xform */
    int32_t has_pre;
    int32_t has_post;
    int32_t has_color;
    int32_t has_speed; /* xform-level color_speed present */
    double pre_A[9];   /* row major, dims x dims */
    double pre_b[3];
    double post_A[9];
    double post_b[3];
    double speed;
    const double *color; /* color_dims values when has_color */
    int32_t n_vars;
    const orc_var *vars;
} orc_xform;

/* a whole flame as written in the JSON, xforms in file order */
typedef struct {
    int32_t dims;
    int32_t color_dims;
    double color_speed;
    uint64_t size[3];
    double lo[3];
    double hi[3];
    int32_t n_xforms;
    const orc_xform *xforms;
    int32_t has_final;
    orc_xform final_xform;
} orc_flame_desc;

/* precision / buffer-word modes */
enum {
    ORC_F64_U64 = 0, /* num_t=double, hist_t=uint64 (reference default build)  */
    ORC_F32_U32 = 1, /* num_t=float,  hist_t=uint32 (reference float build)    */
    ORC_F32_U64 = 2  /* float trajectories, u64 counts, double colour sums     */
};

typedef struct orc_flame orc_flame;

typedef struct {
    uint64_t s_iter;
    uint64_t s_plot;
    uint64_t n_bad;
    int32_t success;           /* 0 when the bad-value limit stopped the render */
    double pt_min[3];
    double pt_max[3];
} orc_stats;

/* build the optimised flame (drop zero weights, sort, cumulative weights) */
orc_flame *orc_flame_new(const orc_flame_desc *desc, int32_t mode,
                         int32_t settle_iters /* -1: 53 / 24 */, char *err,
                         size_t errlen);
void orc_flame_free(orc_flame *f);

/* introspection used by tests */
int32_t orc_flame_n_xforms(const orc_flame *f);
int32_t orc_flame_xform_id(const orc_flame *f, int32_t i);
double orc_flame_cum_weight(const orc_flame *f, int32_t i);
size_t orc_flame_cells(const orc_flame *f);
size_t orc_flame_cell_words(const orc_flame *f);
size_t orc_flame_word_bytes(const orc_flame *f);
double orc_flame_mult_d(const orc_flame *f, int32_t i);

/*
 * Render `samples` samples into `buf` (cells * cell_words words of
 * word_bytes).  Walkers are walker_base, walker_base+1, ...; each renders
 * `batch` samples except the last, which renders the rest.  threads <= 1 runs
 * the walkers one after another on the calling thread (the reference's
 * renderSeeded); threads > 1 runs a std::thread-style pool claiming walkers
 * under a mutex and adding into the shared buffer with relaxed atomics (the
 * reference's render).  xf_dist has one slot per xform in the file.
 * bv_xfs/bv_pts (may be NULL) receive up to bv_cap bad-value records.
 */
int orc_render(orc_flame *f, void *buf, uint64_t seed, uint64_t walker_base,
               uint64_t samples, uint64_t batch, uint64_t bv_limit,
               int32_t threads, uint64_t *xf_dist, orc_stats *stats,
               uint64_t *bv_xfs, double *bv_pts, uint64_t bv_cap);

/*
 * ffr-img's image of a 2-D buffer (src/ffr_img.cpp:199-309): mode 0 mono
 * (8-bit, 255 where the count is not 0), 1 gray, 2 colour (cell_words 4);
 * arithmetic in double (word_bytes 8) or float (4) with libm's log / pow.
 * hist[] receives the count bounds, scaler[] the log(1 + count) bounds.
 * out may be NULL (bounds only).  Returns 0, 1 for an empty histogram
 * (nothing written), 2 for an argument the reference rejects.
 */
int orc_tonemap(const void *buf, uint64_t cells, int32_t cell_words, int32_t word_bytes,
                int32_t mode, double gamma, int32_t bits, void *out, uint64_t hist[2],
                double scaler[2]);

/* the walker stream: word n of walker w under seed */
uint64_t orc_stream_word(uint64_t seed, uint64_t walker, uint64_t n);
/* raw Philox4x32-10 block */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* ISAAC with 2^4-word state, restated from the reference */
typedef struct {
    uint64_t cnt;
    uint64_t rsl[16];
    uint64_t mem[16];
    uint64_t a, b, c;
} orc_isaac64;
typedef struct {
    uint64_t cnt;
    uint32_t rsl[16];
    uint32_t mem[16];
    uint32_t a, b, c;
} orc_isaac32;
void orc_isaac64_seed(orc_isaac64 *s, uint64_t seed);
uint64_t orc_isaac64_next(orc_isaac64 *s);
void orc_isaac32_seed_u32(orc_isaac32 *s, uint32_t seed);
void orc_isaac32_seed_u64(orc_isaac32 *s, uint64_t seed);
uint32_t orc_isaac32_next(orc_isaac32 *s);

/* XForm::applyIteration on one point (xform = sorted index, -1 = final) */
int orc_apply_xform(orc_flame *f, int32_t xform, const double *in, double *out,
                    uint64_t seed, uint64_t walker, uint64_t *draw_index);

/* variation evaluation on one point (tests of single variations) */
int orc_eval_variation(orc_flame *f, int32_t xform, int32_t var,
                       const double *in, double *out, uint64_t seed,
                       uint64_t walker, uint64_t *draw_index);

/* reference libstdc++ std::sort order of xform weights (tests) */
void orc_sort_order(const double *weights, int32_t n, int32_t mode,
                    int32_t *order);

#ifdef __cplusplus
}
#endif
#endif

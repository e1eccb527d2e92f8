/*
 * CPU oracle for the ffr-buf buffer render -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only to check results against.  See oracle.h for what it restates.
 *
 * Parity status (how each part of this file is pinned; tests named in
 * tests/test_oracle_*.py):
 *   orc_isaac64_* / orc_isaac32_*  pinned bit-exact against the reference's
 *       own src/rng/isaac.hpp compiled by oracle/Makefile into oracle/_ref
 *       (fixture tests/golden/isaac_kat.json).
 *   orc_philox4x32_10              pinned by the Random123 published known
 *       answers and by rocRAND's philox4x32_10 engine (library routine).
 *   orc_stream_word / randNum      closed forms (exact dyadic arithmetic).
 *   orc_sort_order                 pinned against libstdc++ std::sort with the
 *       reference's comparator (flame.hpp:291-293) on tied weights.
 *   the render loop (oracle_impl.h) pinned by closed forms, invariants and
 *       brute force; see the header of oracle_impl.h.
 *   orc_tonemap                    pinned by closed forms of the pixel maps
 *       (counts 0 / max / powers of e, gamma 1 and 2, colour = gray x mean
 *       colour, the pix_scale rounding at exactly 1.0) and the reference's
 *       error conditions (tests/test_oracle_image.py).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11; Random123 constants)           */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/*
 * Walker stream contract.  Word n of walker w under seed s is half (n & 1)
 * of the Philox4x32-10 block with key (lo32 s, hi32 s) and counter
 * (lo32 (n>>1), hi32 (n>>1), lo32 w, hi32 w); half 0 is out0 | out1<<32,
 * half 1 is out2 | out3<<32.
 */
uint64_t orc_stream_word(uint64_t seed, uint64_t walker, uint64_t n)
{
    uint64_t blk = n >> 1;
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)walker,
                       (uint32_t)(walker >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    if (n & 1)
        return (uint64_t)out[2] | ((uint64_t)out[3] << 32);
    return (uint64_t)out[0] | ((uint64_t)out[1] << 32);
}

/* sequential reader over one walker's stream */
typedef struct {
    uint64_t seed, walker, n;
    uint64_t cached;   /* word n when n is odd and the block is loaded */
    int have;
} orc_stream;

static void stream_init(orc_stream *s, uint64_t seed, uint64_t walker)
{
    s->seed = seed;
    s->walker = walker;
    s->n = 0;
    s->have = 0;
}

static uint64_t stream_next(orc_stream *s)
{
    uint64_t w;
    if ((s->n & 1) && s->have) {
        w = s->cached;
        s->have = 0;
    } else {
        uint64_t blk = s->n >> 1;
        uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32),
                           (uint32_t)s->walker, (uint32_t)(s->walker >> 32)};
        uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
        uint32_t out[4];
        orc_philox4x32_10(ctr, key, out);
        if (s->n & 1) {
            w = (uint64_t)out[2] | ((uint64_t)out[3] << 32);
            s->have = 0;
        } else {
            w = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
            s->cached = (uint64_t)out[2] | ((uint64_t)out[3] << 32);
            s->have = 1;
        }
    }
    s->n++;
    return w;
}

/* ------------------------------------------------------------------ */
/* ISAAC, 2^4-word state (reference src/rng/isaac.hpp:77-329)           */
/* ------------------------------------------------------------------ */
#define ISAAC_N 16

static void mix64(uint64_t *a, uint64_t *b, uint64_t *c, uint64_t *d,
                  uint64_t *e, uint64_t *f, uint64_t *g, uint64_t *h)
{
    /* Jenkins' 64-bit mix (isaac.hpp:171-182) */
    *a -= *e; *f ^= *h >> 9;  *h += *a;
    *b -= *f; *g ^= *a << 9;  *a += *b;
    *c -= *g; *h ^= *b >> 23; *b += *c;
    *d -= *h; *a ^= *c << 15; *c += *d;
    *e -= *a; *b ^= *d >> 14; *d += *e;
    *f -= *b; *c ^= *e << 20; *e += *f;
    *g -= *c; *d ^= *f >> 17; *f += *g;
    *h -= *d; *e ^= *g << 14; *g += *h;
}

static void mix32(uint32_t *a, uint32_t *b, uint32_t *c, uint32_t *d,
                  uint32_t *e, uint32_t *f, uint32_t *g, uint32_t *h)
{
    /* Jenkins' 32-bit mix (isaac.hpp:158-169) */
    *a ^= *b << 11; *d += *a; *b += *c;
    *b ^= *c >> 2;  *e += *b; *c += *d;
    *c ^= *d << 8;  *f += *c; *d += *e;
    *d ^= *e >> 16; *g += *d; *e += *f;
    *e ^= *f << 10; *h += *e; *f += *g;
    *f ^= *g >> 4;  *a += *f; *g += *h;
    *g ^= *h << 8;  *b += *g; *h += *a;
    *h ^= *a >> 9;  *c += *h; *a += *b;
}

static void isaac64_gen(orc_isaac64 *s)
{
    uint64_t a = s->a, b = s->b + (++s->c);
    uint64_t *mm = s->mem, *r = s->rsl;
    int half = ISAAC_N / 2;
    for (int i = 0; i < ISAAC_N; ++i) {
        uint64_t mixv;
        switch (i & 3) {
        case 0: mixv = ~(a ^ (a << 21)); break;
        case 1: mixv = a ^ (a >> 5); break;
        case 2: mixv = a ^ (a << 12); break;
        default: mixv = a ^ (a >> 33); break;
        }
        uint64_t x = mm[i];
        a = mixv + mm[(i + half) % ISAAC_N];
        uint64_t y = mm[(x >> 3) & (ISAAC_N - 1)] + a + b;
        mm[i] = y;
        b = mm[(y >> (4 + 3)) & (ISAAC_N - 1)] + x;
        r[i] = b;
    }
    s->a = a;
    s->b = b;
}

static void isaac32_gen(orc_isaac32 *s)
{
    uint32_t a = s->a, b = s->b + (++s->c);
    uint32_t *mm = s->mem, *r = s->rsl;
    int half = ISAAC_N / 2;
    for (int i = 0; i < ISAAC_N; ++i) {
        uint32_t mixv;
        switch (i & 3) {
        case 0: mixv = a ^ (a << 13); break;
        case 1: mixv = a ^ (a >> 6); break;
        case 2: mixv = a ^ (a << 2); break;
        default: mixv = a ^ (a >> 16); break;
        }
        uint32_t x = mm[i];
        a = mixv + mm[(i + half) % ISAAC_N];
        uint32_t y = mm[(x >> 2) & (ISAAC_N - 1)] + a + b;
        mm[i] = y;
        b = mm[(y >> (4 + 2)) & (ISAAC_N - 1)] + x;
        r[i] = b;
    }
    s->a = a;
    s->b = b;
}

/* seeding with no seed array: memory is the golden-ratio schedule alone */
static void isaac64_init(orc_isaac64 *s, uint64_t a0, uint64_t b0, uint64_t c0)
{
    uint64_t v[8];
    for (int i = 0; i < 8; ++i)
        v[i] = 0x9e3779b97f4a7c13ULL;
    memset(s->rsl, 0, sizeof s->rsl);
    s->a = a0;
    s->b = b0;
    s->c = c0;
    for (int k = 0; k < 4; ++k)
        mix64(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
    for (int i = 0; i < ISAAC_N; i += 8) {
        mix64(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
        for (int j = 0; j < 8; ++j)
            s->mem[i + j] = v[j];
    }
    isaac64_gen(s);
    s->cnt = ISAAC_N;
}

static void isaac32_init(orc_isaac32 *s, uint32_t a0, uint32_t b0, uint32_t c0)
{
    uint32_t v[8];
    for (int i = 0; i < 8; ++i)
        v[i] = 0x9e3779b9u;
    memset(s->rsl, 0, sizeof s->rsl);
    s->a = a0;
    s->b = b0;
    s->c = c0;
    for (int k = 0; k < 4; ++k)
        mix32(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
    for (int i = 0; i < ISAAC_N; i += 8) {
        mix32(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
        for (int j = 0; j < 8; ++j)
            s->mem[i + j] = v[j];
    }
    isaac32_gen(s);
    s->cnt = ISAAC_N;
}

void orc_isaac64_seed(orc_isaac64 *s, uint64_t seed)
{
    /* isaac.hpp:267-271 */
    isaac64_init(s, seed, ~seed, seed ^ 11713835213681433683ULL);
}

uint64_t orc_isaac64_next(orc_isaac64 *s)
{
    /* words come out from the top of the result array down (isaac.hpp:321-329) */
    if (s->cnt-- == 0) {
        isaac64_gen(s);
        s->cnt = ISAAC_N - 1;
    }
    return s->rsl[s->cnt];
}

void orc_isaac32_seed_u32(orc_isaac32 *s, uint32_t seed)
{
    isaac32_init(s, seed, ~seed, (seed << 16) | (seed >> 16));
}

void orc_isaac32_seed_u64(orc_isaac32 *s, uint64_t seed)
{
    isaac32_init(s, (uint32_t)seed, (uint32_t)(seed >> 32),
                 (uint32_t)(seed ^ (seed >> 32)));
}

uint32_t orc_isaac32_next(orc_isaac32 *s)
{
    if (s->cnt-- == 0) {
        isaac32_gen(s);
        s->cnt = ISAAC_N - 1;
    }
    return s->rsl[s->cnt];
}

/* ------------------------------------------------------------------ */
/* libstdc++ std::sort restated on an index array, comparator           */
/* "weight[a] > weight[b]" (flame.hpp:291-293).  For <= 16 elements it is */
/* the stable insertion sort; above that the introsort with median-of-3 */
/* pivots and the heap-sort fallback.                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *w;
    int32_t *idx;
} sort_ctx;

static int cmp_gt(const sort_ctx *c, int32_t a, int32_t b)
{
    return c->w[a] > c->w[b];
}

static void swap_i(int32_t *x, int32_t *y)
{
    int32_t t = *x;
    *x = *y;
    *y = t;
}

static void unguarded_linear_insert(sort_ctx *c, long last)
{
    int32_t val = c->idx[last];
    long next = last - 1;
    while (cmp_gt(c, val, c->idx[next])) {
        c->idx[last] = c->idx[next];
        last = next;
        --next;
    }
    c->idx[last] = val;
}

static void insertion_sort(sort_ctx *c, long first, long last)
{
    if (first == last)
        return;
    for (long i = first + 1; i != last; ++i) {
        if (cmp_gt(c, c->idx[i], c->idx[first])) {
            int32_t val = c->idx[i];
            memmove(&c->idx[first + 1], &c->idx[first],
                    (size_t)(i - first) * sizeof(int32_t));
            c->idx[first] = val;
        } else {
            unguarded_linear_insert(c, i);
        }
    }
}

static void push_heap(sort_ctx *c, long first, long hole, long top, int32_t val)
{
    long parent = (hole - 1) / 2;
    while (hole > top && cmp_gt(c, c->idx[first + parent], val)) {
        c->idx[first + hole] = c->idx[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    c->idx[first + hole] = val;
}

static void adjust_heap(sort_ctx *c, long first, long hole, long len,
                        int32_t val)
{
    long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (cmp_gt(c, c->idx[first + child], c->idx[first + child - 1]))
            child--;
        c->idx[first + hole] = c->idx[first + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        c->idx[first + hole] = c->idx[first + child - 1];
        hole = child - 1;
    }
    push_heap(c, first, hole, top, val);
}

static void heap_sort_range(sort_ctx *c, long first, long last)
{
    long len = last - first;
    if (len >= 2) {
        long parent = (len - 2) / 2;
        for (;;) {
            adjust_heap(c, first, parent, len, c->idx[first + parent]);
            if (parent == 0)
                break;
            parent--;
        }
    }
    while (last - first > 1) {
        --last;
        int32_t val = c->idx[last];
        c->idx[last] = c->idx[first];
        adjust_heap(c, first, 0, last - first, val);
    }
}

static void move_median_to_first(sort_ctx *c, long result, long a, long b,
                                 long cc)
{
    int32_t *v = c->idx;
    if (cmp_gt(c, v[a], v[b])) {
        if (cmp_gt(c, v[b], v[cc]))
            swap_i(&v[result], &v[b]);
        else if (cmp_gt(c, v[a], v[cc]))
            swap_i(&v[result], &v[cc]);
        else
            swap_i(&v[result], &v[a]);
    } else if (cmp_gt(c, v[a], v[cc]))
        swap_i(&v[result], &v[a]);
    else if (cmp_gt(c, v[b], v[cc]))
        swap_i(&v[result], &v[cc]);
    else
        swap_i(&v[result], &v[b]);
}

static long unguarded_partition(sort_ctx *c, long first, long last, long pivot)
{
    for (;;) {
        while (cmp_gt(c, c->idx[first], c->idx[pivot]))
            ++first;
        --last;
        while (cmp_gt(c, c->idx[pivot], c->idx[last]))
            --last;
        if (!(first < last))
            return first;
        swap_i(&c->idx[first], &c->idx[last]);
        ++first;
    }
}

static void introsort_loop(sort_ctx *c, long first, long last, long depth)
{
    while (last - first > 16) {
        if (depth == 0) {
            heap_sort_range(c, first, last);
            return;
        }
        --depth;
        long mid = first + (last - first) / 2;
        move_median_to_first(c, first, first + 1, mid, last - 1);
        long cut = unguarded_partition(c, first + 1, last, first);
        introsort_loop(c, cut, last, depth);
        last = cut;
    }
}

static void std_sort_desc(const double *w, int32_t *idx, long n)
{
    sort_ctx c = {w, idx};
    if (n <= 0)
        return;
    long lg = 0;
    while ((1L << (lg + 1)) <= n)
        ++lg;
    introsort_loop(&c, 0, n, 2 * lg);
    if (n > 16) {
        insertion_sort(&c, 0, 16);
        for (long i = 16; i != n; ++i)
            unguarded_linear_insert(&c, i);
    } else {
        insertion_sort(&c, 0, n);
    }
}

void orc_sort_order(const double *weights, int32_t n, int32_t mode,
                    int32_t *order)
{
    double *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) {
        /* the comparison happens in num_t */
        w[i] = (mode == ORC_F64_U64) ? weights[i] : (double)(float)weights[i];
        order[i] = i;
    }
    std_sort_desc(w, order, n);
    free(w);
}

/* ------------------------------------------------------------------ */
/* templated render core, once per reference build                      */
/* ------------------------------------------------------------------ */
#define R double
#define CT uint64_t
#define CR double
#define SFX f64u64
#define R_IS_FLOAT 0
#define NATIVE_F32 0
#include "oracle_impl.h"
#undef R
#undef CT
#undef CR
#undef SFX
#undef R_IS_FLOAT
#undef NATIVE_F32

#define R float
#define CT uint32_t
#define CR float
#define SFX f32u32
#define R_IS_FLOAT 1
#define NATIVE_F32 0
#include "oracle_impl.h"
#undef R
#undef CT
#undef CR
#undef SFX
#undef R_IS_FLOAT
#undef NATIVE_F32

#define R float
#define CT uint64_t
#define CR double
#define SFX f32u64
#define R_IS_FLOAT 1
#define NATIVE_F32 1
#include "oracle_impl.h"
#undef R
#undef CT
#undef CR
#undef SFX
#undef R_IS_FLOAT
#undef NATIVE_F32

/* ------------------------------------------------------------------ */
/* mode dispatch                                                        */
/* ------------------------------------------------------------------ */
struct orc_flame {
    int32_t mode;
    void *impl;
};

orc_flame *orc_flame_new(const orc_flame_desc *desc, int32_t mode,
                         int32_t settle_iters, char *err, size_t errlen)
{
    void *impl = NULL;
    if (err && errlen)
        err[0] = 0;
    switch (mode) {
    case ORC_F64_U64: impl = flame_new_f64u64(desc, settle_iters, err, errlen); break;
    case ORC_F32_U32: impl = flame_new_f32u32(desc, settle_iters, err, errlen); break;
    case ORC_F32_U64: impl = flame_new_f32u64(desc, settle_iters, err, errlen); break;
    default:
        if (err && errlen)
            snprintf(err, errlen, "unknown mode");
        return NULL;
    }
    if (!impl)
        return NULL;
    orc_flame *f = (orc_flame *)malloc(sizeof *f);
    f->mode = mode;
    f->impl = impl;
    return f;
}

void orc_flame_free(orc_flame *f)
{
    if (!f)
        return;
    switch (f->mode) {
    case ORC_F64_U64: flame_free_f64u64(f->impl); break;
    case ORC_F32_U32: flame_free_f32u32(f->impl); break;
    default: flame_free_f32u64(f->impl); break;
    }
    free(f);
}

#define DISPATCH(call, ...)                                      \
    switch (f->mode) {                                           \
    case ORC_F64_U64: return call##_f64u64(f->impl, __VA_ARGS__); \
    case ORC_F32_U32: return call##_f32u32(f->impl, __VA_ARGS__); \
    default: return call##_f32u64(f->impl, __VA_ARGS__);          \
    }

int32_t orc_flame_n_xforms(const orc_flame *f) { DISPATCH(flame_nxf, 0) }
int32_t orc_flame_xform_id(const orc_flame *f, int32_t i) { DISPATCH(flame_xfid, i) }
double orc_flame_cum_weight(const orc_flame *f, int32_t i) { DISPATCH(flame_cw, i) }
size_t orc_flame_cells(const orc_flame *f) { DISPATCH(flame_cells, 0) }
size_t orc_flame_cell_words(const orc_flame *f) { DISPATCH(flame_cellwords, 0) }
double orc_flame_mult_d(const orc_flame *f, int32_t i) { DISPATCH(flame_multd, i) }

size_t orc_flame_word_bytes(const orc_flame *f)
{
    return f->mode == ORC_F32_U32 ? 4 : 8;
}

int orc_render(orc_flame *f, void *buf, uint64_t seed, uint64_t walker_base,
               uint64_t samples, uint64_t batch, uint64_t bv_limit,
               int32_t threads, uint64_t *xf_dist, orc_stats *stats,
               uint64_t *bv_xfs, double *bv_pts, uint64_t bv_cap)
{
    DISPATCH(render, buf, seed, walker_base, samples, batch, bv_limit, threads,
             xf_dist, stats, bv_xfs, bv_pts, bv_cap)
}

int orc_apply_xform(orc_flame *f, int32_t xform, const double *in, double *out,
                    uint64_t seed, uint64_t walker, uint64_t *draw_index)
{
    DISPATCH(applyxf, xform, in, out, seed, walker, draw_index)
}

int orc_eval_variation(orc_flame *f, int32_t xform, int32_t var,
                       const double *in, double *out, uint64_t seed,
                       uint64_t walker, uint64_t *draw_index)
{
    DISPATCH(evalvar, xform, var, in, out, seed, walker, draw_index)
}

/* ------------------------------------------------------------------ */
/* ffr-img's image stage (src/ffr_img.cpp:199-309 render_image, over    */
/* ImageRenderer::getValueBounds / renderGrayImage / renderColorImageRGB */
/* src/renderers/image_renderer.hpp:110-192), cell by cell in num_t.     */
/* ------------------------------------------------------------------ */

/* (pix_t)(v * pix_scale): gcc on x86-64 converts through a signed 32-bit
 * integer (cvttsd2si / cvttss2si) and keeps the low bits; NaN and values out
 * of the int32 range give INT32_MIN there, whose low bits are 0 */
static int32_t to_i32_d(double v)
{
    return (v != v || v >= 2147483648.0 || v < -2147483648.0) ? INT32_MIN : (int32_t)v;
}
static int32_t to_i32_f(float v)
{
    return (v != v || v >= 2147483648.0f || v < -2147483648.0f) ? INT32_MIN : (int32_t)v;
}

#define TONEMAP_BODY(N, CT, CR, LOG, POW, EPS, SCALE8, SCALE16, CONV)                   \
    {                                                                                \
        const CT *b = (const CT *)buf;                                               \
        /* getValueBounds of the count and of log(1 + count) (:223-236) */          \
        CT minh = b[0], maxh = b[0];                                                 \
        N mn = LOG((N)1 + (N)b[0]), mx = mn;                                         \
        for (uint64_t c = 0; c < cells; ++c) {                                       \
            const CT h = b[c * (uint64_t)cw];                                        \
            const N l = LOG((N)1 + (N)h);                                            \
            if (h < minh) minh = h;                                                  \
            if (h > maxh) maxh = h;                                                  \
            if (l < mn) mn = l;                                                      \
            if (l > mx) mx = l;                                                      \
        }                                                                            \
        hist[0] = (uint64_t)minh;                                                    \
        hist[1] = (uint64_t)maxh;                                                    \
        scaler[0] = (double)mn;                                                      \
        scaler[1] = (double)mx;                                                      \
        if (mx < EPS)                                                                \
            return 1;                                                                \
        if (!out)                                                                    \
            return 0;                                                                \
        /* num_t gp = 1.0 / arg_gamma: a double division, then num_t */             \
        const N gp = (N)(1.0 / (double)(N)gamma);                                    \
        const N scale = mode == 0 || bits == 8 ? SCALE8 : SCALE16;                   \
        uint8_t *o8 = (uint8_t *)out;                                                \
        uint16_t *o16 = (uint16_t *)out;                                             \
        const int wide = mode != 0 && bits == 16;                                    \
        for (uint64_t c = 0; c < cells; ++c) {                                       \
            const CT *cell = b + c * (uint64_t)cw;                                   \
            N v[3];                                                                  \
            int nv = 1;                                                              \
            if (mode == 0) {                                                         \
                v[0] = cell[0] != 0 ? (N)1.0 : (N)0.0;                               \
            } else {                                                                 \
                const N l = LOG((N)1 + (N)cell[0]) / mx;                             \
                const N ll = POW(l, gp);                                             \
                if (mode == 1) {                                                     \
                    v[0] = ll;                                                       \
                } else {                                                             \
                    const N h = (N)cell[0];                                          \
                    const CR *col = (const CR *)(cell + 1);                          \
                    for (int i = 0; i < 3; ++i)                                      \
                        v[i] = ll * (col[i] / h);                                    \
                    nv = 3;                                                          \
                }                                                                    \
            }                                                                        \
            for (int i = 0; i < nv; ++i) {                                           \
                const int32_t p = CONV(v[i] * scale);                                \
                if (wide)                                                            \
                    o16[c * (uint64_t)nv + (uint64_t)i] = (uint16_t)p;               \
                else                                                                 \
                    o8[c * (uint64_t)nv + (uint64_t)i] = (uint8_t)p;                 \
            }                                                                        \
        }                                                                            \
        return 0;                                                                    \
    }

/* pix_scale (src/types/constants.hpp:77-91): 2^bits (1 - 2^-mantissa) */
static int tonemap_f64(const void *buf, uint64_t cells, int32_t cw, int32_t mode,
                       double gamma, int32_t bits, void *out, uint64_t hist[2],
                       double scaler[2])
TONEMAP_BODY(double, uint64_t, double, log, pow, 1e-20, 256.0 * (1.0 - 0x1p-52),
             65536.0 * (1.0 - 0x1p-52), to_i32_d)

static int tonemap_f32(const void *buf, uint64_t cells, int32_t cw, int32_t mode,
                       double gamma, int32_t bits, void *out, uint64_t hist[2],
                       double scaler[2])
TONEMAP_BODY(float, uint32_t, float, logf, powf, 1e-10f, 256.0f * (1.0f - 0x1p-23f),
             65536.0f * (1.0f - 0x1p-23f), to_i32_f)

int orc_tonemap(const void *buf, uint64_t cells, int32_t cell_words, int32_t word_bytes,
                int32_t mode, double gamma, int32_t bits, void *out, uint64_t hist[2],
                double scaler[2])
{
    if (cells == 0 || (word_bytes != 4 && word_bytes != 8) || mode < 0 || mode > 2 ||
        (bits != 8 && bits != 16) || (mode == 2 && cell_words != 4))
        return 2;
    if (word_bytes == 8)
        return tonemap_f64(buf, cells, cell_words, mode, gamma, bits, out, hist, scaler);
    return tonemap_f32(buf, cells, cell_words, mode, (double)(float)gamma, bits, out, hist,
                       scaler);
}

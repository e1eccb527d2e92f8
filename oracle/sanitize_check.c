/*
 * Sanitizer driver for the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by `make -C oracle sanitize` twice, with oracle.c compiled in:
 * once under -fsanitize=address,undefined and once under -fsanitize=thread.
 * It renders three flames through orc_render with a pool of threads (the
 * restatement of BufferRenderer::render, src/renderers/buffer_renderer.hpp:
 * 269-338: walkers claimed under a mutex, relaxed atomic adds on the shared
 * buffer, CAS loops for colour words, a mutex around the bad-value records),
 * in all three precision modes, and checks that the pooled render equals the
 * same render on one thread: counts exactly, statistics exactly, colours to
 * the rounding of their addition order.  A sanitizer report makes the
 * program exit non-zero (halt_on_error / exitcode options set by the test).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int failures = 0;

#define CHECK(cond, ...)                                                  \
    do {                                                                  \
        if (!(cond)) {                                                    \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);          \
            fprintf(stderr, __VA_ARGS__);                                 \
            fprintf(stderr, "\n");                                        \
            ++failures;                                                   \
        }                                                                 \
    } while (0)

/* an xform with no pre/post affine in the JSON carries the identity
 * (the JSON defaults, src/types/affine.hpp:45-92) */
static void identity_affines(orc_xform *x)
{
    memset(x->pre_A, 0, sizeof x->pre_A);
    memset(x->post_A, 0, sizeof x->post_A);
    x->pre_A[0] = x->pre_A[3] = 1.0;
    x->post_A[0] = x->post_A[3] = 1.0;
}

static void set_affine(double *A, double *b, double a00, double a01, double a10, double a11,
                       double b0, double b1)
{
    A[0] = a00;
    A[1] = a01;
    A[2] = a10;
    A[3] = a11;
    b[0] = b0;
    b[1] = b1;
}

/* Sierpinski triangle: three halving maps, linear */
static void sierpinski(orc_flame_desc *d, orc_xform *xf, orc_var *v)
{
    memset(d, 0, sizeof *d);
    d->dims = 2;
    d->size[0] = d->size[1] = 64;
    d->lo[0] = d->lo[1] = -1.0;
    d->hi[0] = d->hi[1] = 1.0;
    d->color_speed = 0.5;
    memset(v, 0, sizeof *v);
    v->type = 0;
    v->axis_x = v->axis_y = -1;
    v->weight = 1.0;
    const double bx[3] = {0.0, 0.5, 0.25}, by[3] = {0.0, 0.0, 0.5};
    for (int k = 0; k < 3; ++k) {
        memset(&xf[k], 0, sizeof xf[k]);
        identity_affines(&xf[k]);
        xf[k].weight = 1.0;
        xf[k].has_pre = 1;
        set_affine(xf[k].pre_A, xf[k].pre_b, 0.5, 0, 0, 0.5, bx[k], by[k]);
        xf[k].n_vars = 1;
        xf[k].vars = v;
    }
    d->n_xforms = 3;
    d->xforms = xf;
}

/* a colour flame of transcendental and drawing variations with a final
 * xform: swirl + spherical, sinusoidal + horseshoe, polar + blur (blur draws
 * words, so walkers consume their streams unevenly) */
static void colour_flame(orc_flame_desc *d, orc_xform *xf, orc_var (*v)[2],
                         double (*col)[3], orc_var *fv, double *fcol)
{
    memset(d, 0, sizeof *d);
    d->dims = 2;
    d->color_dims = 3;
    d->color_speed = 0.5;
    d->size[0] = 96;
    d->size[1] = 64;
    d->lo[0] = -3.0;
    d->hi[0] = 3.0;
    d->lo[1] = -2.0;
    d->hi[1] = 2.0;
    const int types[3][2] = {{20, 2}, {1, 21}, {22, 8}};
    const double w[3][2] = {{0.6, 0.4}, {0.7, 0.3}, {0.5, 0.5}};
    const double A[3][6] = {{0.6, -0.3, 0.2, 0.5, 0.1, -0.2},
                            {-0.4, 0.5, 0.3, 0.45, -0.3, 0.4},
                            {0.55, 0.1, -0.2, 0.6, 0.2, 0.1}};
    for (int k = 0; k < 3; ++k) {
        memset(&xf[k], 0, sizeof xf[k]);
        identity_affines(&xf[k]);
        xf[k].weight = 1.0 + 0.5 * k;
        xf[k].has_pre = 1;
        set_affine(xf[k].pre_A, xf[k].pre_b, A[k][0], A[k][1], A[k][2], A[k][3], A[k][4],
                   A[k][5]);
        for (int j = 0; j < 2; ++j) {
            memset(&v[k][j], 0, sizeof v[k][j]);
            v[k][j].type = types[k][j];
            v[k][j].axis_x = v[k][j].axis_y = -1;
            v[k][j].weight = w[k][j];
        }
        xf[k].n_vars = 2;
        xf[k].vars = v[k];
        for (int c = 0; c < 3; ++c)
            col[k][c] = 0.1 + 0.3 * k + 0.2 * c;
        xf[k].has_color = 1;
        xf[k].color = col[k];
    }
    d->n_xforms = 3;
    d->xforms = xf;
    memset(fv, 0, sizeof *fv);
    fv->type = 0;
    fv->axis_x = fv->axis_y = -1;
    fv->weight = 1.0;
    memset(&d->final_xform, 0, sizeof d->final_xform);
    identity_affines(&d->final_xform);
    d->has_final = 1;
    d->final_xform.has_pre = 1;
    set_affine(d->final_xform.pre_A, d->final_xform.pre_b, 0.9, 0.1, -0.1, 0.9, 0, 0);
    d->final_xform.n_vars = 1;
    d->final_xform.vars = fv;
    for (int c = 0; c < 3; ++c)
        fcol[c] = 0.8 - 0.2 * c;
    d->final_xform.has_color = 1;
    d->final_xform.has_speed = 1;
    d->final_xform.speed = 0.3;
    d->final_xform.color = fcol;
}

/* a flame that throws a point beyond the bad-value threshold now and then:
 * the pool's bad-value records (under a mutex) and restarts are exercised */
static void occasional_bad(orc_flame_desc *d, orc_xform *xf, orc_var *v)
{
    sierpinski(d, xf, v);
    xf[2].weight = 0.004;
    set_affine(xf[2].pre_A, xf[2].pre_b, 1e25, 0, 0, 1e25, 0, 0);
}

static double rel_err(double a, double b)
{
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return m > 0 ? fabs(a - b) / m : 0.0;
}

static void run(const char *name, const orc_flame_desc *d, int mode, uint64_t samples,
                uint64_t batch, uint64_t bv_limit)
{
    char err[256];
    orc_flame *f = orc_flame_new(d, mode, 20, err, sizeof err);
    CHECK(f != NULL, "%s mode %d: %s", name, mode, err);
    if (!f)
        return;
    const size_t cells = orc_flame_cells(f), cw = orc_flame_cell_words(f),
                 wb = orc_flame_word_bytes(f);
    void *b1 = calloc(cells * cw, wb), *b8 = calloc(cells * cw, wb);
    uint64_t xf1[8] = {0}, xf8[8] = {0};
    uint64_t bx1[64], bx8[64];
    double bp1[192], bp8[192];
    /* orc_render folds its statistics into these (as the reference's
     * getters accumulate): start from zero and empty extremes */
    orc_stats s1, s8;
    memset(&s1, 0, sizeof s1);
    for (int i = 0; i < 3; ++i) {
        s1.pt_min[i] = INFINITY;
        s1.pt_max[i] = -INFINITY;
    }
    s8 = s1;
    int ok1 = orc_render(f, b1, 0x5EED, 3, samples, batch, bv_limit, 1, xf1, &s1, bx1, bp1, 64);
    int ok8 = orc_render(f, b8, 0x5EED, 3, samples, batch, bv_limit, 6, xf8, &s8, bx8, bp8, 64);
    CHECK(ok1 == ok8, "%s mode %d: return %d vs %d", name, mode, ok1, ok8);
    CHECK(s1.s_iter == s8.s_iter && s1.s_plot == s8.s_plot && s1.n_bad == s8.n_bad,
          "%s mode %d: statistics differ", name, mode);
    for (int i = 0; i < d->n_xforms; ++i)
        CHECK(xf1[i] == xf8[i], "%s mode %d: xform %d count", name, mode, i);
    for (int i = 0; i < d->dims; ++i)
        CHECK(s1.pt_min[i] == s8.pt_min[i] && s1.pt_max[i] == s8.pt_max[i],
              "%s mode %d: extremes", name, mode);
    size_t diff = 0;
    double worst = 0;
    for (size_t c = 0; c < cells; ++c) {
        if (wb == 8) {
            const uint64_t *w1 = (const uint64_t *)b1 + c * cw, *w8 = (const uint64_t *)b8 + c * cw;
            diff += w1[0] != w8[0];
            for (size_t k = 1; k < cw; ++k) {
                double x, y;
                memcpy(&x, &w1[k], 8);
                memcpy(&y, &w8[k], 8);
                double e = rel_err(x, y);
                worst = e > worst ? e : worst;
            }
        } else {
            const uint32_t *w1 = (const uint32_t *)b1 + c * cw, *w8 = (const uint32_t *)b8 + c * cw;
            diff += w1[0] != w8[0];
            for (size_t k = 1; k < cw; ++k) {
                float x, y;
                memcpy(&x, &w1[k], 4);
                memcpy(&y, &w8[k], 4);
                double e = rel_err(x, y);
                worst = e > worst ? e : worst;
            }
        }
    }
    CHECK(diff == 0, "%s mode %d: %zu cells with other counts", name, mode, diff);
    CHECK(worst <= (wb == 8 ? 1e-11 : 1e-4), "%s mode %d: colour rel err %g", name, mode,
          worst);
    printf("%-16s mode %d: %llu iterated, %llu plotted, %llu bad, colour rel %.2e\n", name, mode,
           (unsigned long long)s8.s_iter, (unsigned long long)s8.s_plot,
           (unsigned long long)s8.n_bad, worst);
    free(b1);
    free(b8);
    orc_flame_free(f);
}

int main(void)
{
    orc_flame_desc d;
    orc_xform xf[3];
    orc_var v1;
    orc_var v2[3][2];
    double col[3][3], fcol[3];
    orc_var fv;
    for (int mode = 0; mode < 3; ++mode) {
        sierpinski(&d, xf, &v1);
        run("sierpinski", &d, mode, 200000, 1000, 256);
        colour_flame(&d, xf, v2, col, &fv, fcol);
        run("colour", &d, mode, 150000, 700, 256);
        occasional_bad(&d, xf, &v1);
        run("occasional_bad", &d, mode, 200000, 900, 1000000);
    }
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("sanitize check ok\n");
    return 0;
}

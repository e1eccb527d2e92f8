/*
 * CPU oracle render core -- TEST INFRASTRUCTURE ONLY (included by oracle.c
 * once per reference build: R = num_t, CT = hist_t, CR = colour word).
 *
 * Every expression keeps the operand types of the reference source, so the
 * float build's promotions to double (double literals such as 1.0 or M_PI,
 * unqualified C math functions) are reproduced: the C rules for
 * `float op double` and for `float += double` are the C++ rules.
 *
 * Parity status of the parts below (tests in tests/test_oracle_*.py):
 *   flame_new  (drop zero weights, std::sort order, cumulative weights,
 *       mult_d/mult_i)  pinned by closed forms and the libstdc++ sort check.
 *   xf_apply / affine   pinned by IEEE special cases that only the reference
 *       evaluation order b + ((0 + a0 x0) + a1 x1) passes, and by the
 *       Sierpinski invariants (empty central hole, sum of counts = samples).
 *   render loop (init / settle / iterate / bad values / extremes / index)
 *       pinned by brute force on tiny grids, closed-form walkers (a constant
 *       map), the bad-value re-init schedule and the draw-count contract.
 *   variations: every one of the 98 is pinned at one or more points where
 *       the mathematics fixes the image (closed forms, or for variations that
 *       draw, the words the stream holds), and by the draw count of each
 *       (tests/test_oracle_variations.py).  Between those points the
 *       arithmetic rests on the host libm and is "parity unpinned" (the
 *       reference cannot be built here without stand-in headers).
 */

#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define NM(x) ORC_CAT(x, SFX)

#if R_IS_FLOAT
#define EPS 1e-10f
#define BADTH 1e10f
#define SCALE_DOWN (1.0f - 1.0f / (float)(1 << 23))
#define MAXRECT 1e5f
#define SETTLE_DEFAULT 24
#define SINCOSG(a, s, c) sincosf((float)(a), (s), (c))
#define RANDNUM_OF(w) ((R)((w) >> 40) / (float)(1 << 24))
#else
#define EPS 1e-20
#define BADTH 1e20
#define SCALE_DOWN (1.0 - (double)(1.0f / (float)(1LL << 52)))
#define MAXRECT 1e10
#define SETTLE_DEFAULT 53
#define SINCOSG(a, s, c) sincos((double)(a), (s), (c))
#define RANDNUM_OF(w) ((R)((w) >> 11) / (double)(1LL << 53))
#endif

typedef struct {
    int type;
    int ax, ay;
    R weight;
    R p[24];
} NM(var);

typedef struct {
    int id;
    R weight;
    int has_pre, has_post, has_color;
    R preA[9], preb[3], postA[9], postb[3];
    R speed;
    R color[128];
    int nv;
    NM(var) * vars;
} NM(xf);

typedef struct {
    int dims, r;
    uint64_t size[3];
    R lo[3], hi[3];
    R mult_d[3];
    uint64_t mult_i[3];
    uint64_t cells;
    int nxf, n_ids;
    NM(xf) * xf;
    R *cw;
    int has_final;
    NM(xf) fin;
    int settle;
} NM(flame);

/* ---------------- random numbers (flame_rng.hpp:60-219) ---------------- */

static R NM(randnum)(orc_stream *s)
{
    uint64_t w = stream_next(s);
    return RANDNUM_OF(w);
}

static int NM(randbool)(orc_stream *s) { return (int)(stream_next(s) & 1); }

static void NM(randdir)(int dims, orc_stream *s, R *out)
{
    if (dims == 1) {
        out[0] = (R)copysign(1.0, NM(randnum)(s) - 0.5);
    } else if (dims == 2) {
        R a = (2.0 * M_PI) * NM(randnum)(s);
        R sa, ca;
        SINCOSG(a, &sa, &ca);
        out[0] = ca;
        out[1] = sa;
    } else {
        R u = 2.0 * NM(randnum)(s) - 1.0;
        R t = (2.0 * M_PI) * NM(randnum)(s);
        R r = sqrt(1.0 - u * u);
        R st, ct;
        SINCOSG(t, &st, &ct);
        out[0] = r * ct;
        out[1] = r * st;
        out[2] = u;
    }
}

static R NM(randgauss)(orc_stream *s)
{
    R u1 = NM(randnum)(s);
    R u2 = (2.0 * M_PI) * NM(randnum)(s);
    R r = sqrt(-2.0 * log(u1));
    R sn, cs;
    SINCOSG(u2, &sn, &cs);
    return r * cs;
}

/* ---------------- point helpers (point.hpp) ---------------- */

static R NM(norm2sq)(const R *x, int n)
{
    R ret = x[0] * x[0];
    for (int i = 1; i < n; ++i)
        ret += x[i] * x[i];
    return ret;
}

static R NM(norm2)(const R *x, int n)
{
    if (n == 1)
        return (R)fabs(x[0]);
    return (R)sqrt(NM(norm2sq)(x, n));
}

static R NM(norminf)(const R *x, int n)
{
    R ret = (R)fabs(x[0]);
    if (n == 1)
        return ret;
    for (int i = 1; i < n; ++i) {
        R v = (R)fabs(x[i]);
        ret = (ret < v) ? v : ret;
    }
    return ret;
}

static R NM(normsum_p)(const R *x, int n, R p)
{
    R ret = pow(fabs(x[0]), p);
    for (int i = 1; i < n; ++i)
        ret += pow(fabs(x[i]), p);
    return ret;
}

static R NM(norm_p)(const R *x, int n, R p)
{
    return pow(NM(normsum_p)(x, n, p), 1.0 / p);
}

static int NM(bad_value)(R v) { return fabs(v) > BADTH || isnan(v); }

/* ---------------- variation constructors (variations.hpp) ---------------- */

/* raw JSON parameters -> member values, in num_t, as the constructors do */
static int NM(var_setup)(NM(var) * v, const orc_var *raw, int dims, char *err,
                         size_t errlen)
{
    const double *q = raw->params;
    R *p = v->p;
    v->type = raw->type;
    v->ax = raw->axis_x;
    v->ay = raw->axis_y;
    v->weight = (R)raw->weight;
    memset(p, 0, sizeof v->p);
    switch (raw->type) {
    case 3: /* bent: scales_neg, scales_pos */
    case 11: /* separation: params (squared), inside */
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[i];
            p[3 + i] = (R)q[dims + i];
        }
        if (raw->type == 11)
            for (int i = 0; i < dims; ++i)
                p[i] *= p[i];
        break;
    case 4: /* rectangles */
    case 12: /* splits */
        for (int i = 0; i < dims; ++i)
            p[i] = (R)q[i];
        break;
    case 5: /* fisheye */
    case 6: /* bubble */
        p[0] = (R)q[0];
        break;
    case 14: /* modulus: params*2, 1/(params*2) */
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[i];
            p[i] *= (R)2.0;
            p[3 + i] = 1.0 / p[i];
        }
        break;
    case 15: /* celln: sizes, 1/sizes */
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[i];
            p[3 + i] = 1.0 / p[i];
        }
        break;
    case 16: /* spherical_p */
    case 18: /* unit_sphere_p */
        p[0] = (R)q[0];
        if (p[0] <= 0.0) {
            snprintf(err, errlen, "norm <= 0");
            return -1;
        }
        break;
    case 27: { /* disc2: rotation, twist */
        R rot = (R)q[0];
        R twist = (R)q[1];
        p[0] = rot * M_PI;
        R sinadd, cosadd;
        SINCOSG(twist, &sinadd, &cosadd);
        cosadd -= 1.0;
        R k = 1.0 + twist;
        if (twist > 2.0 * M_PI)
            k -= 2.0 * M_PI;
        if (twist < -2.0 * M_PI)
            k += 2.0 * M_PI;
        cosadd *= k;
        sinadd *= k;
        p[1] = cosadd;
        p[2] = sinadd;
        break;
    }
    case 28: /* waves: xfreq xscale yfreq yscale */
        for (int i = 0; i < 4; ++i)
            p[i] = (R)q[i];
        break;
    case 29: { /* fan: x, y */
        R x = (R)q[0];
        R y = (R)q[1];
        p[0] = M_PI * (x * x + EPS);
        p[1] = y;
        break;
    }
    case 30: { /* rings: value */
        R val = (R)q[0];
        p[0] = val * val + EPS;
        break;
    }
    case 39: { /* blob: low high waves */
        R low = (R)q[0];
        R high = (R)q[1];
        p[0] = (high + low) / 2.0;
        p[1] = (high - low) / 2.0;
        p[2] = (R)q[2];
        break;
    }
    case 40: /* pdj: a b c d */
        for (int i = 0; i < 4; ++i)
            p[i] = (R)q[i];
        break;
    case 42: { /* perspective: distance angle */
        p[0] = (R)q[0];
        R angle = (R)q[1];
        p[1] = sin(angle);
        p[2] = p[0] * cos(angle);
        break;
    }
    case 43: /* julian */
    case 44: { /* juliascope: power dist */
        R power = (R)q[0];
        R dist = (R)q[1];
        p[0] = fabs(power);
        p[1] = 1.0 / power;
        p[2] = dist / (2.0 * power);
        break;
    }
    case 45: { /* radial_blur: angle flam3_weight */
        R angle = (R)q[0];
        SINCOSG(angle * M_PI_2, &p[0], &p[1]);
        p[2] = (R)q[1];
        break;
    }
    case 46: /* pie: slices rotation thickness */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        p[3] = (2.0 * M_PI) / p[0];
        break;
    case 47: { /* ngon: sides power corners circle */
        R sides = (R)q[0];
        p[0] = q[1] / 2.0;
        p[1] = (2.0 * M_PI) / sides;
        p[2] = (R)q[2];
        p[3] = (R)q[3];
        p[4] = sides / (2.0 * M_PI);
        break;
    }
    case 48: /* curl: c1 c2 */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        break;
    case 49: /* arch */
    case 51: /* rays */
    case 52: /* blade */
    case 53: /* secant */
    case 54: /* twintrian */
    case 73: /* scry */
    case 95: /* loonie */
        p[0] = (R)q[0];
        if (raw->type == 95)
            p[1] = p[0] * p[0];
        break;
    case 70: /* auger: freq flam3_weight scale sym */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = q[2] / 2.0;
        p[3] = (R)q[3];
        break;
    case 71: /* flux: spread flam3_weight */
        p[0] = 2.0 + q[0];
        p[1] = (R)q[1];
        break;
    case 72: /* mobius: a b c d (2-D points) */
        for (int i = 0; i < 8; ++i)
            p[i] = (R)q[i];
        break;
    case 74: /* split: xsize ysize */
        p[0] = q[0] * M_PI;
        p[1] = q[1] * M_PI;
        break;
    case 75: /* stripes: space warp */
        p[0] = 1.0 - q[0];
        p[1] = (R)q[1];
        break;
    case 76: /* wedge: swirl count angle hole */
    case 78: /* wedge_sph: angle count swirl hole */
    {
        R swirl, count, angle, hole;
        if (raw->type == 76) {
            swirl = (R)q[0]; count = (R)q[1]; angle = (R)q[2]; hole = (R)q[3];
        } else {
            angle = (R)q[0]; count = (R)q[1]; swirl = (R)q[2]; hole = (R)q[3];
        }
        p[0] = swirl;
        p[1] = count;
        p[2] = angle;
        p[3] = hole;
        p[4] = 1.0 - angle * count * (M_1_PI * 0.5);
        break;
    }
    case 77: { /* wedge_julia: angle count power dist */
        R angle = (R)q[0];
        R count = (R)q[1];
        R power = (R)q[2];
        R dist = (R)q[3];
        p[0] = angle;
        p[1] = count;
        p[2] = 1.0 / power;
        p[3] = dist / (2.0 * power);
        p[4] = fabs(power);
        p[5] = 1.0 - angle * count * (M_1_PI * 0.5);
        break;
    }
    case 79: /* whorl: inside outside flam3_weight */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        break;
    case 80: { /* supershape: n1 m n2 n3 rnd holes */
        R n1 = (R)q[0];
        p[0] = q[1] / 4.0;
        p[1] = -1.0 / n1;
        p[2] = (R)q[2];
        p[3] = (R)q[3];
        p[4] = (R)q[4];
        p[5] = (R)q[5];
        break;
    }
    case 81: /* flower: petals holes */
    case 82: /* conic: eccen holes */
    case 83: /* parabola: height width */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        break;
    case 84: /* bipolar: shift */
        p[0] = -M_PI_2 * q[0];
        break;
    case 85: /* boarders: prob */
        p[0] = (R)q[0];
        if (p[0] < 0.0 || p[0] > 1.0) {
            snprintf(err, errlen, "boarders probability out of range");
            return -1;
        }
        break;
    case 87: /* cell: size */
        p[0] = (R)q[0];
        p[1] = 1.0 / p[0];
        break;
    case 88: { /* cpow: r i power */
        R rr = (R)q[0];
        R ii = (R)q[1];
        R power = (R)q[2];
        p[0] = 2.0 * M_PI / power;
        p[1] = rr / power;
        p[2] = ii / power;
        p[3] = power;
        break;
    }
    case 89: { /* curve: xamp yamp xlen ylen */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        R xlen = (R)q[2];
        R ylen = (R)q[3];
        xlen *= xlen;
        ylen *= ylen;
        R e = EPS;
        p[2] = 1.0 / ((e < xlen) ? xlen : e);
        p[3] = 1.0 / ((e < ylen) ? ylen : e);
        break;
    }
    case 92: { /* escher: beta */
        R beta = (R)q[0];
        R seb, ceb;
        SINCOSG(beta, &seb, &ceb);
        p[0] = 0.5 * (1.0 + ceb);
        p[1] = 0.5 * seb;
        break;
    }
    case 94: /* lazysusan: x y spin twist space flam3_weight */
        for (int i = 0; i < 6; ++i)
            p[i] = (R)q[i];
        break;
    case 96: { /* oscope: frequency amplitude damping separation */
        R freq = (R)q[0];
        p[0] = 2.0 * M_PI * freq;
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        p[3] = (R)q[3];
        break;
    }
    case 97: /* popcorn: x y c */
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        break;
    default:
        break;
    }
    return 0;
}

/* ---------------- 2-D variations (variations.hpp:505-2376) ---------------- */

/* LIT(v): a double literal of the reference.  In both reference builds it
 * stays double and promotes float operands (src/types/types.hpp:29 sets
 * num_t; the literals are double).  FP32 trajectories with the default
 * build's 8-byte buffer words have no reference counterpart: there the six
 * basic variations and the colour blend compute in float throughout, and
 * LIT rounds the literal to num_t. */
#undef LIT
#if NATIVE_F32
#define LIT(v) ((R)(v))
#else
#define LIT(v) (v)
#endif

static void NM(calc2d)(const NM(var) * v, R x, R y, R *ox, R *oy,
                       orc_stream *s)
{
    const R *p = v->p;
    R t2[2] = {x, y};
    R n2sq = x * x + y * y;            /* Point<T,2>::norm2sq */
#define N2 ((R)sqrt(n2sq))                /* Point<T,2>::norm2 */
#define ANG ((R)atan2(y, x))              /* Point<T,2>::angle */
    (void)t2;
    switch (v->type) {
    case 20: { /* swirl */
        R r = n2sq, sr, cr;
        SINCOSG(r, &sr, &cr);
        *ox = x * sr - y * cr;
        *oy = x * cr + y * sr;
        return;
    }
    case 21: { /* horseshoe */
        R r = 1.0 / (N2 + EPS);
        R a = (x - y) * (x + y);
        R b = LIT(2.0) * x * y;
        *ox = a * r;
        *oy = b * r;
        return;
    }
    case 22: { /* polar */
        R a = ANG, r = N2;
        *ox = a * LIT(M_1_PI);
        *oy = r - LIT(1.0);
        return;
    }
    case 23: /* polar2 */
        *ox = ANG;
        *oy = log(n2sq);
        return;
    case 24: { /* handkerchief */
        R a = ANG, r = N2;
        R c0 = sin(a + r), c1 = cos(a - r);
        *ox = c0 * r;
        *oy = c1 * r;
        return;
    }
    case 25: { /* heart */
        R a = ANG, r = N2, sa, ca;
        SINCOSG(r * a, &sa, &ca);
        *ox = sa * r;
        *oy = -ca * r;
        return;
    }
    case 26: { /* disc */
        R a = ANG, r = N2, sr, cr;
        SINCOSG(M_PI * r, &sr, &cr);
        *ox = sr * a;
        *oy = cr * a;
        return;
    }
    case 27: { /* disc2 */
        R t = p[0] * (x + y), st, ct;
        SINCOSG(t, &st, &ct);
        R a = ANG;
        *ox = (ct + p[1]) * a;
        *oy = (st + p[2]) * a;
        return;
    }
    case 28: { /* waves */
        R dx = p[1] * sin(y * p[0]);
        R dy = p[3] * sin(x * p[2]);
        *ox = x + dx;
        *oy = y + dy;
        return;
    }
    case 29: { /* fan */
        R dx = p[0], dy = p[1];
        R dx2 = dx * 0.5;
        R a = ANG;
        R m = copysign(1.0, dx2 - fmod(a + dy, dx));
        a += m * dx2;
        R sa, ca;
        SINCOSG(a, &sa, &ca);
        R r = N2;
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 30: { /* rings */
        R dx = p[0];
        R r = N2, sn = y / r, cs = x / r;
        r = fmod(r + dx, 2.0 * dx) - dx + r * (1.0 - dx);
        *ox = cs * r;
        *oy = sn * r;
        return;
    }
    case 31: { /* spiral */
        R r = N2, sa = y / r, ca = x / r, sr, cr;
        SINCOSG(r, &sr, &cr);
        R r1 = 1.0 / (r + EPS);
        *ox = (ca + sr) * r1;
        *oy = (sa - cr) * r1;
        return;
    }
    case 32: { /* hyperbolic */
        R r = N2, sa = y / r, ca = x / r;
        *ox = sa / (r + EPS);
        *oy = ca * r;
        return;
    }
    case 33: { /* diamond */
        R r = N2, sa = y / r, ca = x / r, sr, cr;
        SINCOSG(r, &sr, &cr);
        *ox = sa * cr;
        *oy = ca * sr;
        return;
    }
    case 34: { /* ex */
        R a = ANG, r = N2;
        R n0 = sin(a + r), n1 = cos(a - r);
        R m0 = n0 * n0 * n0 * r, m1 = n1 * n1 * n1 * r;
        *ox = m0 + m1;
        *oy = m0 - m1;
        return;
    }
    case 35: { /* julia */
        R a = 0.5 * ANG + NM(randbool)(s) * M_PI;
        R sa, ca;
        SINCOSG(a, &sa, &ca);
        R r = N2;
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 36: { /* exponential */
        R dx = exp(x - 1.0), sdy, cdy;
        SINCOSG(M_PI * y, &sdy, &cdy);
        *ox = cdy * dx;
        *oy = sdy * dx;
        return;
    }
    case 37: { /* power */
        R r = N2, sa = y / r, ca = x / r;
        R k = pow(r, sa);
        *ox = ca * k;
        *oy = sa * k;
        return;
    }
    case 38: { /* cosine */
        R sa, ca;
        SINCOSG(x * M_PI, &sa, &ca);
        *ox = ca * cosh(y);
        *oy = -sa * sinh(y);
        return;
    }
    case 39: { /* blob */
        R r = N2, sa = y / r, ca = x / r;
        R a = ANG;
        r *= p[0] + p[1] * sin(p[2] * a);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 40: { /* pdj */
        R nx1 = cos(p[1] * x), nx2 = sin(p[2] * x);
        R ny1 = sin(p[0] * y), ny2 = cos(p[3] * y);
        *ox = ny1 - nx1;
        *oy = nx2 - ny2;
        return;
    }
    case 41: /* cylinder */
        *ox = sin(x);
        *oy = y;
        return;
    case 42: { /* perspective */
        R t = 1.0 / (p[0] - y * p[1]);
        R a = p[0] * x, b = p[2] * y;
        *ox = a * t;
        *oy = b * t;
        return;
    }
    case 43: { /* julian */
        int32_t t = (int32_t)trunc(p[0] * NM(randnum)(s));
        R a = (ANG + (2.0 * M_PI) * t) * p[1];
        R r = pow(n2sq, p[2]), sa, ca;
        SINCOSG(a, &sa, &ca);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 44: { /* juliascope */
        int32_t t = (int32_t)trunc(p[0] * NM(randnum)(s));
        R dir;
        NM(randdir)(1, s, &dir);
        R a = ((2.0 * M_PI) * t + dir * ANG) * p[1];
        R r = pow(n2sq, p[2]), sa, ca;
        SINCOSG(a, &sa, &ca);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 45: { /* radial_blur */
        R g = p[2] * NM(randgauss)(s);
        R ra = N2;
        R a = ANG + p[0] * g, sa, ca;
        SINCOSG(a, &sa, &ca);
        R rz = p[1] * g - 1.0;
        *ox = ca * ra + x * rz;
        *oy = sa * ra + y * rz;
        return;
    }
    case 46: { /* pie */
        int32_t sl = (int32_t)(NM(randnum)(s) * p[0] + 0.5);
        R a = p[1] + (sl + NM(randnum)(s) * p[2]) * p[3];
        R r = NM(randnum)(s), sa, ca;
        SINCOSG(a, &sa, &ca);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 47: { /* ngon */
        R r = pow(n2sq, p[0]);
        R theta = ANG;
        R phi = theta - p[1] * floor(theta * p[4]);
        static const R mult[2] = {0.0, 1.0};
        phi -= mult[phi > p[1] * 0.5] * p[1];
        R amp = p[2] * (1.0 / (cos(phi) + EPS) - 1.0) + p[3];
        amp /= r + EPS;
        *ox = x * amp;
        *oy = y * amp;
        return;
    }
    case 48: { /* curl */
        R re = 1.0 + p[0] * x + p[1] * (x * x - y * y);
        R im = p[0] * y + 2.0 * p[1] * x * y;
        R r = 1.0 / (re * re + im * im + EPS);
        R a = x * re + y * im, b = y * re - x * im;
        *ox = a * r;
        *oy = b * r;
        return;
    }
    case 49: { /* arch */
        R a = p[0] * NM(randnum)(s) * M_PI, sa, ca;
        SINCOSG(a, &sa, &ca);
        *ox = sa;
        *oy = sa * sa / ca;
        return;
    }
    case 50: /* tangent */
        *ox = sin(x) / cos(y);
        *oy = tan(y);
        return;
    case 51: { /* rays */
        R a = p[0] * NM(randnum)(s) * M_PI;
        R r = p[0] / (n2sq + EPS);
        R tr = tan(a) * r;
        R c0 = cos(x), c1 = sin(y);
        *ox = c0 * tr;
        *oy = c1 * tr;
        return;
    }
    case 52: { /* blade */
        R r = NM(randnum)(s) * p[0] * N2, sr, cr;
        SINCOSG(r, &sr, &cr);
        *ox = (cr + sr) * x;
        *oy = (cr - sr) * x;
        return;
    }
    case 53: { /* secant */
        R cr = cos(p[0] * N2);
        R icr = 1.0 / cr;
        R sign = copysign(1.0, -cr);
        *ox = x;
        *oy = icr + sign;
        return;
    }
    case 54: { /* twintrian */
        R r = NM(randnum)(s) * p[0] * N2, sr, cr;
        SINCOSG(r, &sr, &cr);
        R diff = log10(sr * sr) + cr;
        if (NM(bad_value)(diff))
            diff = -30.0;
        R a = diff, b = diff - sr * M_PI;
        *ox = a * x;
        *oy = b * x;
        return;
    }
    case 55: { /* cross */
        R q = x * x - y * y;
        R r = sqrt(1.0 / (q * q + EPS));
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 56: { /* exp */
        R e = exp(x), es, ec;
        SINCOSG(y, &es, &ec);
        *ox = ec * e;
        *oy = es * e;
        return;
    }
    case 57: /* log */
        *ox = log(n2sq);
        *oy = ANG;
        return;
    case 58: { /* sin */
        R sn, cs;
        SINCOSG(x, &sn, &cs);
        R sh = sinh(y), ch = cosh(y);
        *ox = sn * ch;
        *oy = cs * sh;
        return;
    }
    case 59: { /* cos */
        R sn, cs;
        SINCOSG(x, &sn, &cs);
        R ch = cosh(y), sh = sinh(y);
        *ox = cs * ch;
        *oy = -sn * sh;
        return;
    }
    case 60: { /* tan */
        R sn, cs;
        SINCOSG(2.0 * x, &sn, &cs);
        R sh = sinh(2.0 * y), ch = cosh(2.0 * y);
        R k = cs + ch;
        R ik = 1 / k;
        *ox = sn * ik;
        *oy = sh * ik;
        return;
    }
    case 61: { /* sec */
        R sn, cs;
        SINCOSG(x, &sn, &cs);
        R sh = sinh(y), ch = cosh(y);
        R a = cs * ch, b = sn * sh;
        R k = cos(2.0 * x) + cosh(2.0 * y);
        R ik = 1 / k;
        *ox = a * ik;
        *oy = b * ik;
        return;
    }
    case 62: { /* csc */
        R sn, cs;
        SINCOSG(x, &sn, &cs);
        R sh = sinh(y), ch = cosh(y);
        R a = sn * ch, b = -cs * sh;
        R k = cosh(2.0 * y) - cos(2.0 * x);
        R ik = 1 / k;
        *ox = a * ik;
        *oy = b * ik;
        return;
    }
    case 63: { /* cot */
        R sn, cs;
        SINCOSG(2.0 * x, &sn, &cs);
        R sh = sinh(2.0 * y), ch = cosh(2.0 * y);
        R k = ch - cs;
        R ik = 1 / k;
        *ox = sn * ik;
        *oy = -sh * ik;
        return;
    }
    case 64: { /* sinh */
        R sn, cs;
        SINCOSG(y, &sn, &cs);
        R sh = sinh(x), ch = cosh(x);
        *ox = sh * cs;
        *oy = ch * sn;
        return;
    }
    case 65: { /* cosh */
        R sn, cs;
        SINCOSG(y, &sn, &cs);
        R sh = sinh(x), ch = cosh(x);
        *ox = ch * cs;
        *oy = sh * sn;
        return;
    }
    case 66: { /* tanh */
        R sn, cs;
        SINCOSG(2.0 * y, &sn, &cs);
        R sh = sinh(2.0 * x), ch = cosh(2.0 * x);
        R k = cs + ch;
        R ik = 1 / k;
        *ox = sh * ik;
        *oy = sn * ik;
        return;
    }
    case 67: { /* sech */
        R sn, cs;
        SINCOSG(y, &sn, &cs);
        R sh = sinh(x), ch = cosh(x);
        R a = cs * ch, b = -sn * sh;
        R k = cos(2.0 * y) + cosh(2.0 * x);
        R ik = 1 / k;
        *ox = a * ik;
        *oy = b * ik;
        return;
    }
    case 68: { /* csch */
        R sn, cs;
        SINCOSG(y, &sn, &cs);
        R sh = sinh(x), ch = cosh(x);
        R a = sh * cs, b = -ch * sn;
        R k = cosh(2.0 * x) - cos(2.0 * y);
        R ik = 1 / k;
        *ox = a * ik;
        *oy = b * ik;
        return;
    }
    case 69: { /* coth */
        R sn, cs;
        SINCOSG(2.0 * y, &sn, &cs);
        R sh = sinh(2.0 * x), ch = cosh(2.0 * x);
        R k = ch - cs;
        R ik = 1 / k;
        *ox = sh * ik;
        *oy = sn * ik;
        return;
    }
    case 70: { /* auger */
        R sx = sin(p[0] * x);
        R ty = sin(p[0] * y);
        R dy = y + p[1] * (p[2] + fabs(y)) * sx;
        R dx = x + p[1] * (p[2] + fabs(x)) * ty;
        *ox = x + p[3] * (dx - x);
        *oy = dy;
        return;
    }
    case 71: { /* flux */
        R xpw = x + p[1];
        R xmw = x - p[1];
        R y2 = y * y;
        R avgr = p[0] * sqrt(sqrt(y2 + xpw * xpw) / sqrt(y2 + xmw * xmw));
        R avga = (atan2(y, xmw) - atan2(y, xpw)) * 0.5;
        R c, sn;
        SINCOSG(avga, &c, &sn); /* sine into the first, as the reference */
        *ox = c * avgr;
        *oy = sn * avgr;
        return;
    }
    case 72: { /* mobius */
        R re_u = p[0] * x - p[1] * y + p[2];
        R im_u = p[0] * y + p[1] * x + p[3];
        R re_v = p[4] * x - p[5] * y + p[6];
        R im_v = p[4] * y + p[5] * x + p[7];
        R rad = 1.0 / (re_v * re_v + im_v * im_v + EPS);
        R a = re_u * re_v + im_u * im_v, b = im_u * re_v - re_u * im_v;
        *ox = a * rad;
        *oy = b * rad;
        return;
    }
    case 73: { /* scry */
        R t = n2sq;
        R r = 1.0 / (sqrt(t) * (t + 1.0 / (p[0] + EPS)));
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 74: { /* split */
        R xs = copysign(1.0, cos(x * p[0]));
        R ys = copysign(1.0, cos(y * p[1]));
        *ox = x * ys;
        *oy = y * xs;
        return;
    }
    case 75: { /* stripes */
        R rx = floor(x + 0.5);
        R ox0 = x - rx;
        *ox = ox0 * p[0] + rx;
        *oy = y + ox0 * ox0 * p[1];
        return;
    }
    case 76: { /* wedge */
        R r = N2;
        R a = ANG + p[0] * r;
        R c = floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        a = a * p[4] + c * p[2];
        R sa, ca;
        SINCOSG(a, &sa, &ca);
        R k = r + p[3];
        *ox = ca * k;
        *oy = sa * k;
        return;
    }
    case 77: { /* wedge_julia */
        R r = pow(n2sq, p[3]);
        int32_t tr = (int32_t)(p[4] * NM(randnum)(s));
        R a = (ANG + (2.0 * M_PI) * tr) * p[2];
        R c = floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        R sa, ca;
        a = a * p[5] + c * p[0];
        SINCOSG(a, &sa, &ca);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 78: { /* wedge_sph */
        R r = 1.0 / (N2 + EPS);
        R a = ANG + p[0] * r;
        R c = floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        R sa, ca;
        a = a * p[4] + c * p[2];
        SINCOSG(a, &sa, &ca);
        R k = r + p[3];
        *ox = ca * k;
        *oy = sa * k;
        return;
    }
    case 79: { /* whorl */
        R r = N2;
        R a = ANG;
        a += p[r >= p[2]] / (p[2] - r);
        R sa, ca;
        SINCOSG(a, &sa, &ca);
        *ox = ca * r;
        *oy = sa * r;
        return;
    }
    case 80: { /* supershape */
        R theta = p[0] * ANG + M_PI_4;
        R st, ct;
        SINCOSG(theta, &st, &ct);
        R t1 = pow(fabs(ct), p[2]);
        R t2 = pow(fabs(st), p[3]);
        R tr = N2;
        R r = (p[4] * NM(randnum)(s) + (1.0 - p[4]) * tr) - p[5];
        r *= pow(t1 + t2, p[1]) / tr;
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 81: { /* flower */
        R theta = ANG;
        R r = (NM(randnum)(s) - p[1]) * cos(p[0] * theta);
        r /= N2 + EPS;
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 82: { /* conic */
        R tr = N2;
        R ct = x / (tr + EPS);
        R r = (NM(randnum)(s) - p[1]) * p[0] / (tr + tr * p[0] * ct);
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 83: { /* parabola */
        R sr, cr;
        SINCOSG(N2, &sr, &cr);
        R a = p[0] * sr * sr * NM(randnum)(s);
        R b = p[1] * cr * NM(randnum)(s);
        *ox = a;
        *oy = b;
        return;
    }
    case 84: { /* bipolar */
        R x2y2 = n2sq;
        R t = x2y2 + 1.0;
        R x2 = 2.0 * x;
        R yy = 0.5 * atan2(2.0 * y, x2y2 - 1.0) + p[0];
        yy -= M_PI * floor(yy * M_1_PI + 0.5);
        *ox = log((t + x2) / (t - x2));
        *oy = yy;
        return;
    }
    case 85: { /* boarders */
        R rx = rint(x), ry = rint(y);
        R ox0 = x - rx, oy0 = y - ry;
        if (NM(randnum)(s) >= p[0]) {
            *ox = ox0 * 0.5 + rx;
            *oy = oy0 * 0.5 + ry;
        } else {
            R mag = 1.0 - p[0];
            if (fabs(ox0) >= fabs(oy0)) {
                R sg = copysign(mag, ox0);
                *ox = ox0 * 0.5 + rx + sg;
                *oy = oy0 * 0.5 + ry + sg * oy0 / ox0;
            } else {
                R sg = copysign(mag, oy0);
                *ox = ox0 * 0.5 + rx + sg * ox0 / oy0;
                *oy = oy0 * 0.5 + ry + sg;
            }
        }
        return;
    }
    case 86: { /* butterfly */
        R y2 = 2.0 * y;
        R r = sqrt(fabs(x * y) / (x * x + y2 * y2 + EPS));
        *ox = x * r;
        *oy = y2 * r;
        return;
    }
    case 87: { /* cell */
        R cx = floor(x * p[1]);
        R cy = floor(y * p[1]);
        R dx = x - cx * p[0];
        R dy = y - cy * p[0];
        R xs = copysign(2.0, cx);
        R ys = copysign(2.0, cy);
        R x2 = cx * xs;
        R y2 = cy * ys;
        x2 -= (R)(cx < 0);
        y2 -= (R)(cy < 0);
        *ox = dx + x2 * p[0];
        *oy = -dy - y2 * p[0];
        return;
    }
    case 88: { /* cpow */
        R a = ANG;
        R lnr = 0.5 * log(n2sq);
        R ang = p[1] * a + p[2] * lnr + p[0] * floor(p[3] * NM(randnum)(s));
        R sa, ca;
        SINCOSG(ang, &sa, &ca);
        R k = exp(p[1] * lnr - p[2] * a);
        *ox = ca * k;
        *oy = sa * k;
        return;
    }
    case 89: { /* curve */
        R vx = p[0] * exp(-y * y * p[2]);
        R vy = p[1] * exp(-x * x * p[3]);
        *ox = x + vx;
        *oy = y + vy;
        return;
    }
    case 90: { /* edisc */
        R tmp = n2sq + 1.0;
        R tmp2 = 2.0 * x;
        R xmax = 0.5 * (sqrt(tmp + tmp2) + sqrt(tmp - tmp2));
        R a1 = log(xmax + sqrt(xmax - 1.0));
        R a2 = -acos(x / xmax);
        R s1, c1;
        SINCOSG(a1, &s1, &c1);
        R s2 = sinh(a2), c2 = cosh(a2);
        s1 *= copysign(1.0, -y);
        *ox = c2 * c1;
        *oy = s2 * s1;
        return;
    }
    case 91: { /* elliptic */
        R tmp = n2sq + 1.0;
        R x2 = 2.0 * x;
        R xmax = 0.5 * (sqrt(tmp + x2) + sqrt(tmp - x2));
        R a = x / xmax;
        R b = 1.0 - a * a;
        R ssx = xmax - 1.0;
        b = b < 0.0 ? 0.0 : sqrt(b);
        ssx = ssx < 0.0 ? 0.0 : sqrt(ssx);
        *ox = atan2(a, b);
        *oy = copysign(1.0, y) * log(xmax + ssx);
        return;
    }
    case 92: { /* escher */
        R a = ANG;
        R lnr = 0.5 * log(n2sq);
        R n = p[0] * a + p[1] * lnr;
        R sn, cn;
        SINCOSG(n, &sn, &cn);
        R k = exp(p[0] * lnr - p[1] * a);
        *ox = cn * k;
        *oy = sn * k;
        return;
    }
    case 93: { /* foci */
        R expx = 0.5 * exp(x);
        R expnx = 0.25 / expx;
        R sn, cn;
        SINCOSG(y, &sn, &cn);
        R tmp = 1.0 / (expx + expnx - cn);
        R a = expx - expnx;
        *ox = a * tmp;
        *oy = sn * tmp;
        return;
    }
    case 94: { /* lazysusan */
        R lx = x - p[0];
        R ly = y + p[1];
        R r = hypot(lx, ly);
        if (r < p[5]) {
            R a = atan2(ly, lx) + p[2] + p[3] * (p[5] - r);
            R sa, ca;
            SINCOSG(a, &sa, &ca);
            *ox = r * ca + p[0];
            *oy = r * sa - p[1];
        } else {
            r = 1.0 + p[4] / (r + EPS);
            *ox = r * lx + p[0];
            *oy = r * ly - p[1];
        }
        return;
    }
    case 95: { /* loonie */
        R r2 = n2sq;
        R w2 = p[1];
        R r = p[0];
        if (r2 < w2)
            r *= sqrt(w2 / (r2 + EPS) - 1.0);
        *ox = x * r;
        *oy = y * r;
        return;
    }
    case 96: { /* oscope */
        R damp = exp(-fabs(x) * p[2]);
        R t = p[1] * damp * cos(p[0] * x) + p[3];
        R yy = copysign(1.0, fabs(y) - t) * y;
        *ox = x;
        *oy = yy;
        return;
    }
    case 97: { /* popcorn */
        R dx = p[0] * sin(tan(y * p[2]));
        R dy = p[1] * sin(tan(x * p[2]));
        *ox = x + dx;
        *oy = y + dy;
        return;
    }
    default:
        *ox = 0;
        *oy = 0;
        return;
    }
#undef N2
#undef ANG
}

/* ---------------- N-D variations and the 2-D embedding ---------------- */

static void NM(calc)(const NM(var) * v, int dims, const R *t, R *out,
                     orc_stream *s)
{
    const R *p = v->p;
    int n = dims;
    if (v->type >= 20) {
        /* VariationFrom2D (variations.hpp:63-107) */
        int ax = (dims == 2) ? 0 : v->ax;
        int ay = (dims == 2) ? 1 : v->ay;
        R ox, oy;
        NM(calc2d)(v, t[ax], t[ay], &ox, &oy, s);
        for (int i = 0; i < dims; ++i)
            out[i] = 0;
        out[ax] = ox;
        out[ay] = oy;
        return;
    }
    switch (v->type) {
    case 0: /* linear */
        for (int i = 0; i < n; ++i)
            out[i] = t[i];
        return;
    case 1: /* sinusoidal */
        for (int i = 0; i < n; ++i)
            out[i] = sin(t[i]);
        return;
    case 2: { /* spherical */
        R r = 1.0 / (NM(norm2sq)(t, n) + EPS);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 3: /* bent */
        for (int i = 0; i < n; ++i) {
            R c = t[i];
            if (c < 0.0)
                c *= p[i];
            else
                c *= p[3 + i];
            out[i] = c;
        }
        return;
    case 4: /* rectangles */
        for (int i = 0; i < n; ++i) {
            R q = p[i], x = t[i];
            if (q == 0.0)
                out[i] = x;
            else
                out[i] = (2.0 * floor(x / q) + 1.0) * q - x;
        }
        return;
    case 5: { /* fisheye */
        R r = 1.0 / (NM(norm2)(t, n) + p[0]);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 6: { /* bubble */
        R r = 1.0 / (NM(norm2sq)(t, n) + p[0]);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 7: { /* noise */
        R r = NM(randnum)(s);
        R dir[3];
        NM(randdir)(n, s, dir);
        for (int i = 0; i < n; ++i) {
            R m = t[i] * dir[i];
            out[i] = m * r;
        }
        return;
    }
    case 8: { /* blur */
        R r = NM(randnum)(s);
        R dir[3];
        NM(randdir)(n, s, dir);
        for (int i = 0; i < n; ++i)
            out[i] = dir[i] * r;
        return;
    }
    case 9: /* gaussian_blur */
    case 13: { /* pre_blur */
        R g = NM(randgauss)(s);
        R dir[3];
        NM(randdir)(n, s, dir);
        for (int i = 0; i < n; ++i)
            out[i] = dir[i] * g;
        return;
    }
    case 10: /* square_noise */
        for (int i = 0; i < n; ++i)
            out[i] = NM(randnum)(s) - 0.5;
        return;
    case 11: /* separation */
        for (int i = 0; i < n; ++i) {
            R sg = copysign(1.0, t[i]);
            out[i] = sg * (sqrt(t[i] * t[i] + p[i]) - sg * p[3 + i]);
        }
        return;
    case 12: /* splits */
        for (int i = 0; i < n; ++i) {
            R sg = copysign(1.0, t[i]);
            R q = sg * p[i];
            out[i] = t[i] + q;
        }
        return;
    case 14: /* modulus */
        for (int i = 0; i < n; ++i)
            out[i] = t[i] - p[i] * floor(t[i] * p[3 + i] + 0.5);
        return;
    case 15: /* celln */
        for (int i = 0; i < n; ++i) {
            R x = floor(t[i] * p[3 + i]);
            R dx = t[i] - x * p[i];
            R xs = copysign(2.0, x);
            R x2 = x * xs;
            x2 -= (R)(x < 0);
            out[i] = dx + x2 * p[i];
        }
        return;
    case 16: { /* spherical_p */
        R r = 1.0 / (NM(normsum_p)(t, n, p[0]) + EPS);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 17: { /* unit_sphere */
        R r = 1.0 / (NM(norm2)(t, n) + EPS);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 18: { /* unit_sphere_p */
        R r = 1.0 / (NM(norm_p)(t, n, p[0]) + EPS);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    case 19: { /* unit_cube */
        R r = 1.0 / (NM(norminf)(t, n) + EPS);
        for (int i = 0; i < n; ++i)
            out[i] = t[i] * r;
        return;
    }
    default:
        for (int i = 0; i < n; ++i)
            out[i] = 0;
        return;
    }
}

/* ---------------- affine and xform (affine.hpp:104-110, xform.hpp:211-227) */

static void NM(affine)(const R *A, const R *b, int n, const R *x, R *out)
{
    for (int i = 0; i < n; ++i) {
        R dot = 0;
        for (int j = 0; j < n; ++j)
            dot += A[i * n + j] * x[j];
        out[i] = b[i] + dot;
    }
}

static void NM(xf_apply)(const NM(xf) * xf, int dims, R *pt, orc_stream *s)
{
    R t[3], v[3] = {0, 0, 0}, vv[3];
    if (dims < 3 || xf->has_pre)
        NM(affine)(xf->preA, xf->preb, dims, pt, t);
    else
        for (int i = 0; i < dims; ++i)
            t[i] = pt[i];
    for (int k = 0; k < xf->nv; ++k) {
        NM(calc)(&xf->vars[k], dims, t, vv, s);
        R w = xf->vars[k].weight;
        for (int i = 0; i < dims; ++i)
            v[i] += vv[i] * w;
    }
    if (dims < 3 || xf->has_post)
        NM(affine)(xf->postA, xf->postb, dims, v, pt);
    else
        for (int i = 0; i < dims; ++i)
            pt[i] = v[i];
}

/* ---------------- flame setup (flame.hpp:44-83, buffer_renderer:114-140) */

static int NM(xf_setup)(NM(xf) * o, const orc_xform *in, int dims, int r,
                        R default_speed, int id, char *err, size_t errlen)
{
    memset(o, 0, sizeof *o);
    o->id = id;
    o->weight = (R)in->weight;
    o->has_pre = in->has_pre;
    o->has_post = in->has_post;
    for (int i = 0; i < dims * dims; ++i) {
        o->preA[i] = (R)in->pre_A[i];
        o->postA[i] = (R)in->post_A[i];
    }
    for (int i = 0; i < dims; ++i) {
        o->preb[i] = (R)in->pre_b[i];
        o->postb[i] = (R)in->post_b[i];
    }
    o->has_color = (r > 0 && in->has_color) ? 1 : 0;
    if (o->has_color)
        for (int i = 0; i < r; ++i)
            o->color[i] = (R)in->color[i];
    o->speed = (r > 0 && in->has_speed) ? (R)in->speed : default_speed;
    o->vars = (NM(var) *)calloc((size_t)(in->n_vars > 0 ? in->n_vars : 1),
                                sizeof(NM(var)));
    o->nv = 0;
    for (int k = 0; k < in->n_vars; ++k) {
        NM(var) tmp;
        if (NM(var_setup)(&tmp, &in->vars[k], dims, err, errlen) != 0)
            return -1;
        /* XForm::_optimize drops variations of weight 0 (xform.hpp:47-58) */
        if (fabs(tmp.weight) == 0.0)
            continue;
        o->vars[o->nv++] = tmp;
    }
    return 0;
}

static void NM(flame_free)(void *vf)
{
    NM(flame) *f = (NM(flame) *)vf;
    if (!f)
        return;
    for (int i = 0; i < f->nxf; ++i)
        free(f->xf[i].vars);
    free(f->xf);
    free(f->cw);
    if (f->has_final)
        free(f->fin.vars);
    free(f);
}

static void *NM(flame_new)(const orc_flame_desc *d, int32_t settle, char *err,
                           size_t errlen)
{
    NM(flame) *f = (NM(flame) *)calloc(1, sizeof *f);
    f->dims = d->dims;
    f->r = d->color_dims;
    f->settle = settle >= 0 ? settle : SETTLE_DEFAULT;
    R default_speed = (R)d->color_speed;
    int n = d->n_xforms;
    NM(xf) *all = (NM(xf) *)calloc((size_t)(n > 0 ? n : 1), sizeof(NM(xf)));
    for (int i = 0; i < n; ++i) {
        if (NM(xf_setup)(&all[i], &d->xforms[i], f->dims, f->r, default_speed,
                         i, err, errlen) != 0) {
            for (int j = 0; j <= i; ++j)
                free(all[j].vars);
            free(all);
            free(f);
            return NULL;
        }
    }
    f->n_ids = n;
    if (d->has_final) {
        if (NM(xf_setup)(&f->fin, &d->final_xform, f->dims, f->r,
                         default_speed, -1, err, errlen) != 0) {
            for (int j = 0; j < n; ++j)
                free(all[j].vars);
            free(all);
            free(f);
            return NULL;
        }
        f->has_final = 1;
    }
    /* Flame::_optimize: drop weight-0 xforms (compared in num_t) */
    double *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int32_t *keep = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (all[i].weight == 0.0) {
            free(all[i].vars);
            continue;
        }
        keep[m++] = i;
    }
    if (m == 0) {
        snprintf(err, errlen, "Flame(): no xforms remaining after optimization");
        free(w);
        free(keep);
        free(all);
        if (f->has_final)
            free(f->fin.vars);
        free(f);
        return NULL;
    }
    for (int i = 0; i < m; ++i)
        w[i] = (double)all[keep[i]].weight;
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    for (int i = 0; i < m; ++i)
        order[i] = i;
    std_sort_desc(w, order, m);
    f->nxf = m;
    f->xf = (NM(xf) *)calloc((size_t)m, sizeof(NM(xf)));
    for (int i = 0; i < m; ++i)
        f->xf[i] = all[keep[order[i]]];
    free(order);
    free(keep);
    free(w);
    free(all);
    /* Flame::_setupCumulativeWeights */
    f->cw = (R *)calloc((size_t)m, sizeof(R));
    R normdiv = 0.0;
    for (int i = 0; i < m; ++i)
        normdiv += f->xf[i].weight;
    R wsum = 0.0;
    for (int i = 0; i < m; ++i) {
        wsum += f->xf[i].weight / normdiv;
        f->cw[i] = wsum;
    }
    f->cw[m - 1] = 1.0;
    /* BufferRenderer::_init */
    f->cells = 1;
    for (int i = 0; i < f->dims; ++i) {
        f->size[i] = d->size[i];
        f->lo[i] = (R)d->lo[i];
        f->hi[i] = (R)d->hi[i];
        f->mult_d[i] = (R)(f->size[i]) / (f->hi[i] - f->lo[i]);
        f->mult_d[i] *= SCALE_DOWN;
        f->mult_i[i] = f->cells;
        f->cells *= f->size[i];
        if (f->cells >= (1ULL << 48)) {
            snprintf(err, errlen, "BufferRenderer(): histogram too big");
            NM(flame_free)(f);
            return NULL;
        }
    }
    return f;
}

static int32_t NM(flame_nxf)(void *vf, int unused)
{
    (void)unused;
    return ((NM(flame) *)vf)->nxf;
}
static int32_t NM(flame_xfid)(void *vf, int32_t i)
{
    return ((NM(flame) *)vf)->xf[i].id;
}
static double NM(flame_cw)(void *vf, int32_t i)
{
    return (double)((NM(flame) *)vf)->cw[i];
}
static size_t NM(flame_cells)(void *vf, int unused)
{
    (void)unused;
    return ((NM(flame) *)vf)->cells;
}
static size_t NM(flame_cellwords)(void *vf, int unused)
{
    (void)unused;
    return 1 + (size_t)((NM(flame) *)vf)->r;
}
static double NM(flame_multd)(void *vf, int32_t i)
{
    return (double)((NM(flame) *)vf)->mult_d[i];
}

/* ---------------- one walker (render_iterator.hpp, buffer_renderer:150-250) */

typedef struct {
    pthread_mutex_t mu;
    uint64_t n_bad;
    uint64_t *bv_xfs;
    double *bv_pts;
    uint64_t bv_cap;
    int atomic;
} NM(shared);

static const NM(xf) * NM(pick)(const NM(flame) * f, orc_stream *s)
{
    int i = 0;
    R r = NM(randnum)(s);
    while (f->cw[i] < r)
        ++i;
    return &f->xf[i];
}

static void NM(walker_init)(const NM(flame) * f, R *p, R *c, orc_stream *s)
{
    for (int i = 0; i < f->dims; ++i)
        p[i] = 2.0 * NM(randnum)(s) - 1.0;
    for (int k = 0; k < f->settle; ++k)
        NM(xf_apply)(NM(pick)(f, s), f->dims, p, s);
    for (int i = 0; i < f->r; ++i)
        c[i] = NM(randnum)(s);
}

static void NM(add_color)(CR *dst, CR v, int atomic)
{
    if (!atomic) {
        *dst += v;
        return;
    }
    CR old;
    __atomic_load(dst, &old, __ATOMIC_RELAXED);
    CR nv;
    do {
        nv = old + v;
    } while (!__atomic_compare_exchange(dst, &old, &nv, 1, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED));
}

/* returns 1 when the walker finished its samples, 0 when stopped */
static int NM(walker)(const NM(flame) * f, CT *buf, uint64_t seed,
                      uint64_t walker, uint64_t samples, uint64_t bv_limit,
                      NM(shared) * sh, uint64_t *xf_dist, uint64_t *s_iter_out,
                      uint64_t *s_plot_out, double *pmin, double *pmax)
{
    const int dims = f->dims, r = f->r;
    const size_t cw = 1 + (size_t)r;
    if (__atomic_load_n(&sh->n_bad, __ATOMIC_RELAXED) > bv_limit)
        return 0;
    orc_stream st;
    stream_init(&st, seed, walker);
    R p[3] = {0, 0, 0}, pf[3] = {0, 0, 0};
    R c[128], cf[128];
    memset(c, 0, sizeof c);
    memset(cf, 0, sizeof cf);
    NM(walker_init)(f, p, c, &st);
    R s_iter = 0, s_plot = 0;
    R mn[3], mx[3];
    for (int i = 0; i < dims; ++i) {
        mn[i] = (R)pmin[i];
        mx[i] = (R)pmax[i];
    }
    for (; samples > 0; --samples) {
        /* RenderIterator::iterate */
        const NM(xf) *xf = NM(pick)(f, &st);
        R sp = xf->speed;
        int xf_id = xf->id;
        NM(xf_apply)(xf, dims, p, &st);
        if (xf->has_color)
            for (int i = 0; i < r; ++i)
                c[i] = (LIT(1.0) - sp) * c[i] + sp * xf->color[i];
        if (f->has_final) {
            const NM(xf) *ff = &f->fin;
            sp = ff->speed;
            for (int i = 0; i < dims; ++i)
                pf[i] = p[i];
            NM(xf_apply)(ff, dims, pf, &st);
            if (ff->has_color)
                for (int i = 0; i < r; ++i)
                    cf[i] = (LIT(1.0) - sp) * c[i] + sp * ff->color[i];
            else
                for (int i = 0; i < r; ++i)
                    cf[i] = c[i];
        } else {
            for (int i = 0; i < dims; ++i)
                pf[i] = p[i];
            for (int i = 0; i < r; ++i)
                cf[i] = c[i];
        }
        ++s_iter;
        ++xf_dist[xf_id];
        /* bad value: record, then re-init (pf and cf are not refreshed) */
        int bad = 0;
        for (int i = 0; i < dims; ++i)
            if (NM(bad_value)(p[i]))
                bad = 1;
        if (bad) {
            uint64_t nb;
            pthread_mutex_lock(&sh->mu);
            nb = sh->n_bad;
            if (nb < sh->bv_cap) {
                if (sh->bv_xfs)
                    sh->bv_xfs[nb] = (uint64_t)xf_id;
                if (sh->bv_pts)
                    for (int i = 0; i < dims; ++i)
                        sh->bv_pts[nb * 3 + i] = (double)p[i];
            }
            /* other walkers read the count without the lock (the limit
             * test at their start): the store is atomic */
            nb = sh->n_bad + 1;
            __atomic_store_n(&sh->n_bad, nb, __ATOMIC_RELAXED);
            pthread_mutex_unlock(&sh->mu);
            if (nb > bv_limit)
                break;
            NM(walker_init)(f, p, c, &st);
        }
        for (int i = 0; i < dims; ++i) {
            if (p[i] < mn[i])
                mn[i] = p[i];
            if (p[i] > mx[i])
                mx[i] = p[i];
        }
        /* in bounds (NaN counts as out of bounds) */
        int in = 1;
        for (int i = 0; i < dims; ++i)
            if (!(pf[i] >= f->lo[i] && pf[i] <= f->hi[i]))
                in = 0;
        if (!in)
            continue;
        ++s_plot;
        uint64_t bi = (uint64_t)((pf[0] - f->lo[0]) * f->mult_d[0]);
        for (int i = 1; i < dims; ++i) {
            uint64_t di = (uint64_t)((pf[i] - f->lo[i]) * f->mult_d[i]);
            bi += di * f->mult_i[i];
        }
        if (bi >= f->cells)
            continue; /* the reference would write past its buffer */
        CT *cell = buf + bi * cw;
        if (sh->atomic)
            __atomic_fetch_add(cell, (CT)1, __ATOMIC_RELAXED);
        else
            *cell += 1;
        CR *col = (CR *)(cell + 1);
        for (int i = 0; i < r; ++i)
            NM(add_color)(&col[i], (CR)cf[i], sh->atomic);
    }
    *s_iter_out += (uint64_t)s_iter;
    *s_plot_out += (uint64_t)s_plot;
    for (int i = 0; i < dims; ++i) {
        pmin[i] = (double)mn[i];
        pmax[i] = (double)mx[i];
    }
    return samples == 0;
}

/* ---------------- the pool (buffer_renderer.hpp:269-372) ---------------- */

typedef struct {
    const NM(flame) * f;
    CT *buf;
    uint64_t seed, walker_base, samples, batch, bv_limit;
    NM(shared) * sh;
    pthread_mutex_t mu;
    uint64_t claimed;
    int success;
    uint64_t s_iter, s_plot;
    uint64_t *xf_dist;
    double pmin[3], pmax[3];
} NM(pool);

static void *NM(pool_thread)(void *arg)
{
    NM(pool) *P = (NM(pool) *)arg;
    const NM(flame) *f = P->f;
    uint64_t *xf_loc = (uint64_t *)calloc((size_t)(f->n_ids > 0 ? f->n_ids : 1),
                                          sizeof(uint64_t));
    uint64_t s_iter = 0, s_plot = 0;
    for (;;) {
        pthread_mutex_lock(&P->mu);
        uint64_t left = P->samples - P->claimed;
        uint64_t n = left < P->batch ? left : P->batch;
        uint64_t walker = P->walker_base + P->claimed / P->batch;
        P->claimed += n;
        double mn[3], mx[3];
        for (int i = 0; i < 3; ++i) {
            mn[i] = P->pmin[i];
            mx[i] = P->pmax[i];
        }
        pthread_mutex_unlock(&P->mu);
        if (n == 0)
            break;
        int ok = NM(walker)(f, P->buf, P->seed, walker, n, P->bv_limit, P->sh,
                            xf_loc, &s_iter, &s_plot, mn, mx);
        pthread_mutex_lock(&P->mu);
        for (int i = 0; i < f->dims; ++i) {
            if (mn[i] < P->pmin[i])
                P->pmin[i] = mn[i];
            if (mx[i] > P->pmax[i])
                P->pmax[i] = mx[i];
        }
        if (!ok)
            P->success = 0;
        pthread_mutex_unlock(&P->mu);
        if (!ok)
            break;
    }
    pthread_mutex_lock(&P->mu);
    P->s_iter += s_iter;
    P->s_plot += s_plot;
    for (int i = 0; i < f->n_ids; ++i)
        P->xf_dist[i] += xf_loc[i];
    pthread_mutex_unlock(&P->mu);
    free(xf_loc);
    return NULL;
}

static int NM(render)(void *vf, void *buf, uint64_t seed, uint64_t walker_base,
                      uint64_t samples, uint64_t batch, uint64_t bv_limit,
                      int32_t threads, uint64_t *xf_dist, orc_stats *stats,
                      uint64_t *bv_xfs, double *bv_pts, uint64_t bv_cap)
{
    NM(flame) *f = (NM(flame) *)vf;
    NM(shared) sh;
    pthread_mutex_init(&sh.mu, NULL);
    sh.n_bad = 0;
    sh.bv_xfs = bv_xfs;
    sh.bv_pts = bv_pts;
    sh.bv_cap = bv_cap;
    sh.atomic = threads > 1;
    NM(pool) P;
    memset(&P, 0, sizeof P);
    P.f = f;
    P.buf = (CT *)buf;
    P.seed = seed;
    P.walker_base = walker_base;
    P.samples = samples;
    P.batch = batch ? batch : 1;
    P.bv_limit = bv_limit;
    P.sh = &sh;
    pthread_mutex_init(&P.mu, NULL);
    P.success = 1;
    P.xf_dist = xf_dist;
    for (int i = 0; i < 3; ++i) {
        P.pmin[i] = stats->pt_min[i];
        P.pmax[i] = stats->pt_max[i];
    }
    if (samples > 0) {
        if (threads <= 1) {
            NM(pool_thread)(&P);
        } else {
            uint64_t nb = (samples + P.batch - 1) / P.batch;
            int nt = threads;
            if ((uint64_t)nt > nb)
                nt = (int)nb;
            pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
            for (int i = 0; i < nt; ++i)
                pthread_create(&th[i], NULL, NM(pool_thread), &P);
            for (int i = 0; i < nt; ++i)
                pthread_join(th[i], NULL);
            free(th);
        }
    }
    stats->s_iter += P.s_iter;
    stats->s_plot += P.s_plot;
    stats->n_bad = sh.n_bad;
    stats->success = P.success;
    for (int i = 0; i < 3; ++i) {
        stats->pt_min[i] = P.pmin[i];
        stats->pt_max[i] = P.pmax[i];
    }
    pthread_mutex_destroy(&P.mu);
    pthread_mutex_destroy(&sh.mu);
    return P.success ? 0 : 1;
}

static int NM(applyxf)(void *vf, int32_t xform, const double *in, double *out,
                       uint64_t seed, uint64_t walker, uint64_t *draw_index)
{
    NM(flame) *f = (NM(flame) *)vf;
    if (xform < -1 || xform >= f->nxf || (xform == -1 && !f->has_final))
        return -1;
    const NM(xf) *xf = (xform < 0) ? &f->fin : &f->xf[xform];
    orc_stream st;
    stream_init(&st, seed, walker);
    st.n = *draw_index;
    R p[3];
    for (int i = 0; i < f->dims; ++i)
        p[i] = (R)in[i];
    NM(xf_apply)(xf, f->dims, p, &st);
    for (int i = 0; i < f->dims; ++i)
        out[i] = (double)p[i];
    *draw_index = st.n;
    return 0;
}

static int NM(evalvar)(void *vf, int32_t xform, int32_t var, const double *in,
                       double *out, uint64_t seed, uint64_t walker,
                       uint64_t *draw_index)
{
    NM(flame) *f = (NM(flame) *)vf;
    const NM(xf) *xf = (xform < 0) ? &f->fin : &f->xf[xform];
    if (var < 0 || var >= xf->nv)
        return -1;
    orc_stream st;
    stream_init(&st, seed, walker);
    st.n = *draw_index;
    R t[3], o[3];
    for (int i = 0; i < f->dims; ++i)
        t[i] = (R)in[i];
    NM(calc)(&xf->vars[var], f->dims, t, o, &st);
    for (int i = 0; i < f->dims; ++i)
        out[i] = (double)o[i];
    *draw_index = st.n;
    return 0;
}

#undef EPS
#undef BADTH
#undef SCALE_DOWN
#undef MAXRECT
#undef SETTLE_DEFAULT
#undef SINCOSG
#undef RANDNUM_OF
#undef NM
#undef ORC_CAT
#undef ORC_CAT2

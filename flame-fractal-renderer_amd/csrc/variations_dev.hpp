// Device variation functions (reference: src/variations/variations.hpp).
//
// var_eval<R, D, VS, NATIVE>(type, E, t, out, st) evaluates variation `type` on the
// pre-affine point t.  E points at the variation's table entry:
// [weight, (axis_x, axis_y when a 2-D variation runs in 3-D), params...]
// where params are the values the reference constructor precomputes
// (computed on the host in num_t, flame_model.cpp var_setup).  Expressions
// keep the reference's operand types: double literals promote float
// operands to double exactly as in the reference's float build (NATIVE:
// FP32 trajectories with 8-byte buffer words compute the six basic
// variations in float, see native_f32).
// The caller multiplies by the weight and sums; random words are drawn from
// the walker stream in the reference's order.
//
// VS selects which types a kernel carries, so linear-only and six-variation
// flames get small kernels: 0 = linear only, 1 = the six variations of the
// benchmark flames (linear, sinusoidal, spherical, swirl, horseshoe, polar),
// 2 = all 98.
#pragma once

#include "device_common.hpp"

namespace ffr {

enum { VS_LINEAR = 0, VS_BASIC = 1, VS_ALL = 2 };

__host__ __device__ constexpr bool vs_has(int vs, int type)
{
    return vs == VS_ALL || type == 0 ||
           (vs == VS_BASIC && (type == 1 || type == 2 || type == 20 || type == 21 || type == 22));
}

template <typename R> __device__ __forceinline__ R dev_randnum(Stream &s)
{
    return rand_num(s, R(0));
}

// randDirection<dims> (flame_rng.hpp:171-206)
template <typename R, int D> __device__ __forceinline__ void dev_randdir(Stream &s, R *out)
{
    if constexpr (D == 1) {
        out[0] = (R)m_copysign(1.0, dev_randnum<R>(s) - 0.5);
    } else if constexpr (D == 2) {
        R a = (2.0 * M_PI) * dev_randnum<R>(s);
        R sa, ca;
        m_sincos(a, &sa, &ca);
        out[0] = ca;
        out[1] = sa;
    } else {
        R u = 2.0 * dev_randnum<R>(s) - 1.0;
        R t = (2.0 * M_PI) * dev_randnum<R>(s);
        R r = m_sqrt(1.0 - u * u);
        R st, ct;
        m_sincos(t, &st, &ct);
        out[0] = r * ct;
        out[1] = r * st;
        out[2] = u;
    }
}

// randGaussian via Box-Muller (flame_rng.hpp:115-148)
template <typename R> __device__ __forceinline__ R dev_randgauss(Stream &s)
{
    R u1 = dev_randnum<R>(s);
    R u2 = (2.0 * M_PI) * dev_randnum<R>(s);
    R r = m_sqrt(-2.0 * m_log((double)u1));
    R sn, cs;
    m_sincos(u2, &sn, &cs);
    return r * cs;
}

template <typename R> __device__ __forceinline__ R dev_norm2sq(const R *x, int n)
{
    R ret = x[0] * x[0];
    for (int i = 1; i < n; ++i)
        ret += x[i] * x[i];
    return ret;
}

// The six basic variations of one 2-D point, for kernels that evaluate each
// once per iteration: linear, sinusoidal, spherical, swirl, horseshoe,
// polar (codes 0..5).  Each output is the expression calc2d / var_eval
// evaluate for it (variations.hpp:169-208, 509-555), with the squared norm
// and its root shared.  `need` has bit k set when some lane of the wave
// uses code k; the others are left untouched.
// 1.0 / x as the reference writes it for spherical and horseshoe: with
// num_t = float that is a double division whose quotient is then rounded to
// float.  Double rounding of a quotient is innocuous when the wide format
// has at least 2p + 2 bits (Figueroa 1995; 53 >= 2 * 24 + 2), so the
// correctly rounded float division gives the same bits, at float cost.
template <typename R> __device__ __forceinline__ R recip_ref(R x) { return (R)1 / x; }

template <typename R, bool NATIVE>
__device__ __forceinline__ void basic6(unsigned need, const R *t, R (&o)[6][2])
{
    using L = lit_t<R, NATIVE>;
    const R EPS = Consts<R>::eps;
    const R x = t[0], y = t[1];
    const R n2sq = x * x + y * y;
    o[0][0] = x;
    o[0][1] = y;
    if (need & 2u) {
        o[1][0] = m_sin(x);
        o[1][1] = m_sin(y);
    }
    if (need & 4u) {
        R r = recip_ref(n2sq + EPS);
        o[2][0] = x * r;
        o[2][1] = y * r;
    }
    if (need & 8u) {
        R sr, cr;
        m_sincos(n2sq, &sr, &cr);
        o[3][0] = x * sr - y * cr;
        o[3][1] = x * cr + y * sr;
    }
    R n2 = 0;
    if (need & 48u)
        n2 = (R)m_sqrt(n2sq);
    if (need & 16u) {
        R r = recip_ref(n2 + EPS);
        R a = (x - y) * (x + y);
        R b = (L)2.0 * x * y;
        o[4][0] = a * r;
        o[4][1] = b * r;
    }
    if (need & 32u) {
        R a = m_atan2(y, x);
        o[5][0] = a * (L)M_1_PI;
        o[5][1] = n2 - (L)1.0;
    }
}

// 2-D variations (variations.hpp:505-2376)
template <typename R, int VS, bool NATIVE>
__device__ __forceinline__ void calc2d(int type, const R *p, R x, R y, R &ox, R &oy, Stream &st)
{
    using L = lit_t<R, NATIVE>;
    const R EPS = Consts<R>::eps;
    const R n2sq = x * x + y * y;
#define N2 ((R)m_sqrt(n2sq))
#define ANG (m_atan2(y, x))
    switch (type) {
    case 20: if constexpr (vs_has(VS, 20)) { // swirl
        R r = n2sq, sr, cr;
        m_sincos(r, &sr, &cr);
        ox = x * sr - y * cr;
        oy = x * cr + y * sr;
    } break;
    case 21: if constexpr (vs_has(VS, 21)) { // horseshoe
        R r = recip_ref(N2 + EPS);
        R a = (x - y) * (x + y);
        R b = (L)2.0 * x * y;
        ox = a * r;
        oy = b * r;
    } break;
    case 22: if constexpr (vs_has(VS, 22)) { // polar
        R a = ANG, r = N2;
        ox = a * (L)M_1_PI;
        oy = r - (L)1.0;
    } break;
    case 23: if constexpr (vs_has(VS, 23)) { // polar2
        ox = ANG;
        oy = m_log(n2sq);
    } break;
    case 24: if constexpr (vs_has(VS, 24)) { // handkerchief
        R a = ANG, r = N2;
        R c0 = m_sin(a + r), c1 = m_cos(a - r);
        ox = c0 * r;
        oy = c1 * r;
    } break;
    case 25: if constexpr (vs_has(VS, 25)) { // heart
        R a = ANG, r = N2, sa, ca;
        m_sincos(r * a, &sa, &ca);
        ox = sa * r;
        oy = -ca * r;
    } break;
    case 26: if constexpr (vs_has(VS, 26)) { // disc
        R a = ANG, r = N2, sr, cr;
        m_sincos((R)(M_PI * r), &sr, &cr);
        ox = sr * a;
        oy = cr * a;
    } break;
    case 27: if constexpr (vs_has(VS, 27)) { // disc2
        R t = p[0] * (x + y), sn, cs;
        m_sincos(t, &sn, &cs);
        R a = ANG;
        ox = (cs + p[1]) * a;
        oy = (sn + p[2]) * a;
    } break;
    case 28: if constexpr (vs_has(VS, 28)) { // waves
        R dx = p[1] * m_sin(y * p[0]);
        R dy = p[3] * m_sin(x * p[2]);
        ox = x + dx;
        oy = y + dy;
    } break;
    case 29: if constexpr (vs_has(VS, 29)) { // fan
        R dx = p[0], dy = p[1];
        R dx2 = dx * 0.5;
        R a = ANG;
        R m = m_copysign(1.0, dx2 - m_fmod((double)(a + dy), (double)dx));
        a += m * dx2;
        R sa, ca;
        m_sincos(a, &sa, &ca);
        R r = N2;
        ox = ca * r;
        oy = sa * r;
    } break;
    case 30: if constexpr (vs_has(VS, 30)) { // rings
        R dx = p[0];
        R r = N2, sn = y / r, cs = x / r;
        r = m_fmod((double)(r + dx), 2.0 * dx) - dx + r * (1.0 - dx);
        ox = cs * r;
        oy = sn * r;
    } break;
    case 31: if constexpr (vs_has(VS, 31)) { // spiral
        R r = N2, sa = y / r, ca = x / r, sr, cr;
        m_sincos(r, &sr, &cr);
        R r1 = 1.0 / (r + EPS);
        ox = (ca + sr) * r1;
        oy = (sa - cr) * r1;
    } break;
    case 32: if constexpr (vs_has(VS, 32)) { // hyperbolic
        R r = N2, sa = y / r, ca = x / r;
        ox = sa / (r + EPS);
        oy = ca * r;
    } break;
    case 33: if constexpr (vs_has(VS, 33)) { // diamond
        R r = N2, sa = y / r, ca = x / r, sr, cr;
        m_sincos(r, &sr, &cr);
        ox = sa * cr;
        oy = ca * sr;
    } break;
    case 34: if constexpr (vs_has(VS, 34)) { // ex
        R a = ANG, r = N2;
        R n0 = m_sin(a + r), n1 = m_cos(a - r);
        R m0 = n0 * n0 * n0 * r, m1 = n1 * n1 * n1 * r;
        ox = m0 + m1;
        oy = m0 - m1;
    } break;
    case 35: if constexpr (vs_has(VS, 35)) { // julia
        R a = 0.5 * ANG + (int)(st.next() & 1) * M_PI;
        R sa, ca;
        m_sincos(a, &sa, &ca);
        R r = N2;
        ox = ca * r;
        oy = sa * r;
    } break;
    case 36: if constexpr (vs_has(VS, 36)) { // exponential
        R dx = m_exp(x - 1.0), sdy, cdy;
        m_sincos((R)(M_PI * y), &sdy, &cdy);
        ox = cdy * dx;
        oy = sdy * dx;
    } break;
    case 37: if constexpr (vs_has(VS, 37)) { // power
        R r = N2, sa = y / r, ca = x / r;
        R k = m_pow(r, sa);
        ox = ca * k;
        oy = sa * k;
    } break;
    case 38: if constexpr (vs_has(VS, 38)) { // cosine
        R sa, ca;
        m_sincos((R)(x * M_PI), &sa, &ca);
        ox = ca * m_cosh(y);
        oy = -sa * m_sinh(y);
    } break;
    case 39: if constexpr (vs_has(VS, 39)) { // blob
        R r = N2, sa = y / r, ca = x / r;
        R a = ANG;
        r *= p[0] + p[1] * m_sin(p[2] * a);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 40: if constexpr (vs_has(VS, 40)) { // pdj
        R nx1 = m_cos(p[1] * x), nx2 = m_sin(p[2] * x);
        R ny1 = m_sin(p[0] * y), ny2 = m_cos(p[3] * y);
        ox = ny1 - nx1;
        oy = nx2 - ny2;
    } break;
    case 41: if constexpr (vs_has(VS, 41)) { // cylinder
        ox = m_sin(x);
        oy = y;
    } break;
    case 42: if constexpr (vs_has(VS, 42)) { // perspective
        R t = 1.0 / (p[0] - y * p[1]);
        R a = p[0] * x, b = p[2] * y;
        ox = a * t;
        oy = b * t;
    } break;
    case 43: if constexpr (vs_has(VS, 43)) { // julian
        int32_t t = (int32_t)m_trunc(p[0] * dev_randnum<R>(st));
        R a = (ANG + (2.0 * M_PI) * t) * p[1];
        R r = m_pow(n2sq, p[2]), sa, ca;
        m_sincos(a, &sa, &ca);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 44: if constexpr (vs_has(VS, 44)) { // juliascope
        int32_t t = (int32_t)m_trunc(p[0] * dev_randnum<R>(st));
        R dir;
        dev_randdir<R, 1>(st, &dir);
        R a = ((2.0 * M_PI) * t + dir * ANG) * p[1];
        R r = m_pow(n2sq, p[2]), sa, ca;
        m_sincos(a, &sa, &ca);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 45: if constexpr (vs_has(VS, 45)) { // radial_blur
        R g = p[2] * dev_randgauss<R>(st);
        R ra = N2;
        R a = ANG + p[0] * g, sa, ca;
        m_sincos(a, &sa, &ca);
        R rz = p[1] * g - 1.0;
        ox = ca * ra + x * rz;
        oy = sa * ra + y * rz;
    } break;
    case 46: if constexpr (vs_has(VS, 46)) { // pie
        int32_t sl = (int32_t)(dev_randnum<R>(st) * p[0] + 0.5);
        R a = p[1] + (sl + dev_randnum<R>(st) * p[2]) * p[3];
        R r = dev_randnum<R>(st), sa, ca;
        m_sincos(a, &sa, &ca);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 47: if constexpr (vs_has(VS, 47)) { // ngon
        R r = m_pow(n2sq, p[0]);
        R theta = ANG;
        R phi = theta - p[1] * m_floor((double)(theta * p[4]));
        phi -= (phi > p[1] * 0.5 ? (R)1.0 : (R)0.0) * p[1];
        R amp = p[2] * (1.0 / (m_cos((double)phi) + EPS) - 1.0) + p[3];
        amp /= r + EPS;
        ox = x * amp;
        oy = y * amp;
    } break;
    case 48: if constexpr (vs_has(VS, 48)) { // curl
        R re = 1.0 + p[0] * x + p[1] * (x * x - y * y);
        R im = p[0] * y + 2.0 * p[1] * x * y;
        R r = 1.0 / (re * re + im * im + EPS);
        R a = x * re + y * im, b = y * re - x * im;
        ox = a * r;
        oy = b * r;
    } break;
    case 49: if constexpr (vs_has(VS, 49)) { // arch
        R a = p[0] * dev_randnum<R>(st) * M_PI, sa, ca;
        m_sincos(a, &sa, &ca);
        ox = sa;
        oy = sa * sa / ca;
    } break;
    case 50: if constexpr (vs_has(VS, 50)) { // tangent
        ox = m_sin((double)x) / m_cos((double)y);
        oy = m_tan(y);
    } break;
    case 51: if constexpr (vs_has(VS, 51)) { // rays
        R a = p[0] * dev_randnum<R>(st) * M_PI;
        R r = p[0] / (n2sq + EPS);
        R tr = m_tan((double)a) * r;
        R c0 = m_cos(x), c1 = m_sin(y);
        ox = c0 * tr;
        oy = c1 * tr;
    } break;
    case 52: if constexpr (vs_has(VS, 52)) { // blade
        R r = dev_randnum<R>(st) * p[0] * N2, sr, cr;
        m_sincos(r, &sr, &cr);
        ox = (cr + sr) * x;
        oy = (cr - sr) * x;
    } break;
    case 53: if constexpr (vs_has(VS, 53)) { // secant
        R cr = m_cos(p[0] * N2);
        R icr = 1.0 / cr;
        R sign = m_copysign(1.0, -(double)cr);
        ox = x;
        oy = icr + sign;
    } break;
    case 54: if constexpr (vs_has(VS, 54)) { // twintrian
        R r = dev_randnum<R>(st) * p[0] * N2, sr, cr;
        m_sincos(r, &sr, &cr);
        R diff = m_log10((double)(sr * sr)) + cr;
        if (bad_value(diff))
            diff = -30.0;
        R a = diff, b = diff - sr * M_PI;
        ox = a * x;
        oy = b * x;
    } break;
    case 55: if constexpr (vs_has(VS, 55)) { // cross
        R q = x * x - y * y;
        R r = m_sqrt(1.0 / (q * q + EPS));
        ox = x * r;
        oy = y * r;
    } break;
    case 56: if constexpr (vs_has(VS, 56)) { // exp
        R e = m_exp(x), es, ec;
        m_sincos(y, &es, &ec);
        ox = ec * e;
        oy = es * e;
    } break;
    case 57: if constexpr (vs_has(VS, 57)) { // log
        ox = m_log(n2sq);
        oy = ANG;
    } break;
    case 58: if constexpr (vs_has(VS, 58)) { // sin
        R sn, cs;
        m_sincos(x, &sn, &cs);
        R sh = m_sinh(y), ch = m_cosh(y);
        ox = sn * ch;
        oy = cs * sh;
    } break;
    case 59: if constexpr (vs_has(VS, 59)) { // cos
        R sn, cs;
        m_sincos(x, &sn, &cs);
        R ch = m_cosh(y), sh = m_sinh(y);
        ox = cs * ch;
        oy = -sn * sh;
    } break;
    case 60: if constexpr (vs_has(VS, 60)) { // tan
        R sn, cs;
        m_sincos((R)(2.0 * x), &sn, &cs);
        R sh = m_sinh(2.0 * y), ch = m_cosh(2.0 * y);
        R k = cs + ch;
        R ik = 1 / k;
        ox = sn * ik;
        oy = sh * ik;
    } break;
    case 61: if constexpr (vs_has(VS, 61)) { // sec
        R sn, cs;
        m_sincos(x, &sn, &cs);
        R sh = m_sinh(y), ch = m_cosh(y);
        R a = cs * ch, b = sn * sh;
        R k = m_cos(2.0 * x) + m_cosh(2.0 * y);
        R ik = 1 / k;
        ox = a * ik;
        oy = b * ik;
    } break;
    case 62: if constexpr (vs_has(VS, 62)) { // csc
        R sn, cs;
        m_sincos(x, &sn, &cs);
        R sh = m_sinh(y), ch = m_cosh(y);
        R a = sn * ch, b = -cs * sh;
        R k = m_cosh(2.0 * y) - m_cos(2.0 * x);
        R ik = 1 / k;
        ox = a * ik;
        oy = b * ik;
    } break;
    case 63: if constexpr (vs_has(VS, 63)) { // cot
        R sn, cs;
        m_sincos((R)(2.0 * x), &sn, &cs);
        R sh = m_sinh(2.0 * y), ch = m_cosh(2.0 * y);
        R k = ch - cs;
        R ik = 1 / k;
        ox = sn * ik;
        oy = -sh * ik;
    } break;
    case 64: if constexpr (vs_has(VS, 64)) { // sinh
        R sn, cs;
        m_sincos(y, &sn, &cs);
        R sh = m_sinh(x), ch = m_cosh(x);
        ox = sh * cs;
        oy = ch * sn;
    } break;
    case 65: if constexpr (vs_has(VS, 65)) { // cosh
        R sn, cs;
        m_sincos(y, &sn, &cs);
        R sh = m_sinh(x), ch = m_cosh(x);
        ox = ch * cs;
        oy = sh * sn;
    } break;
    case 66: if constexpr (vs_has(VS, 66)) { // tanh
        R sn, cs;
        m_sincos((R)(2.0 * y), &sn, &cs);
        R sh = m_sinh(2.0 * x), ch = m_cosh(2.0 * x);
        R k = cs + ch;
        R ik = 1 / k;
        ox = sh * ik;
        oy = sn * ik;
    } break;
    case 67: if constexpr (vs_has(VS, 67)) { // sech
        R sn, cs;
        m_sincos(y, &sn, &cs);
        R sh = m_sinh(x), ch = m_cosh(x);
        R a = cs * ch, b = -sn * sh;
        R k = m_cos(2.0 * y) + m_cosh(2.0 * x);
        R ik = 1 / k;
        ox = a * ik;
        oy = b * ik;
    } break;
    case 68: if constexpr (vs_has(VS, 68)) { // csch
        R sn, cs;
        m_sincos(y, &sn, &cs);
        R sh = m_sinh(x), ch = m_cosh(x);
        R a = sh * cs, b = -ch * sn;
        R k = m_cosh(2.0 * x) - m_cos(2.0 * y);
        R ik = 1 / k;
        ox = a * ik;
        oy = b * ik;
    } break;
    case 69: if constexpr (vs_has(VS, 69)) { // coth
        R sn, cs;
        m_sincos((R)(2.0 * y), &sn, &cs);
        R sh = m_sinh(2.0 * x), ch = m_cosh(2.0 * x);
        R k = ch - cs;
        R ik = 1 / k;
        ox = sh * ik;
        oy = sn * ik;
    } break;
    case 70: if constexpr (vs_has(VS, 70)) { // auger
        R sx = m_sin(p[0] * x);
        R ty = m_sin(p[0] * y);
        R dy = y + p[1] * (p[2] + m_fabs((double)y)) * sx;
        R dx = x + p[1] * (p[2] + m_fabs((double)x)) * ty;
        ox = x + p[3] * (dx - x);
        oy = dy;
    } break;
    case 71: if constexpr (vs_has(VS, 71)) { // flux
        R xpw = x + p[1];
        R xmw = x - p[1];
        R y2 = y * y;
        R avgr = p[0] * m_sqrt(m_sqrt((double)(y2 + xpw * xpw)) /
                               m_sqrt((double)(y2 + xmw * xmw)));
        R avga = (m_atan2((double)y, (double)xmw) - m_atan2((double)y, (double)xpw)) * 0.5;
        R c, sn;
        m_sincos(avga, &c, &sn); // sine into the first output, as the reference
        ox = c * avgr;
        oy = sn * avgr;
    } break;
    case 72: if constexpr (vs_has(VS, 72)) { // mobius
        R re_u = p[0] * x - p[1] * y + p[2];
        R im_u = p[0] * y + p[1] * x + p[3];
        R re_v = p[4] * x - p[5] * y + p[6];
        R im_v = p[4] * y + p[5] * x + p[7];
        R rad = 1.0 / (re_v * re_v + im_v * im_v + EPS);
        R a = re_u * re_v + im_u * im_v, b = im_u * re_v - re_u * im_v;
        ox = a * rad;
        oy = b * rad;
    } break;
    case 73: if constexpr (vs_has(VS, 73)) { // scry
        R t = n2sq;
        R r = 1.0 / (m_sqrt((double)t) * (t + 1.0 / (p[0] + EPS)));
        ox = x * r;
        oy = y * r;
    } break;
    case 74: if constexpr (vs_has(VS, 74)) { // split
        R xs = m_copysign(1.0, (double)m_cos(x * p[0]));
        R ys = m_copysign(1.0, (double)m_cos(y * p[1]));
        ox = x * ys;
        oy = y * xs;
    } break;
    case 75: if constexpr (vs_has(VS, 75)) { // stripes
        R rx = m_floor(x + 0.5);
        R ox0 = x - rx;
        ox = ox0 * p[0] + rx;
        oy = y + ox0 * ox0 * p[1];
    } break;
    case 76: if constexpr (vs_has(VS, 76)) { // wedge
        R r = N2;
        R a = ANG + p[0] * r;
        R c = m_floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        a = a * p[4] + c * p[2];
        R sa, ca;
        m_sincos(a, &sa, &ca);
        R k = r + p[3];
        ox = ca * k;
        oy = sa * k;
    } break;
    case 77: if constexpr (vs_has(VS, 77)) { // wedge_julia
        R r = m_pow(n2sq, p[3]);
        int32_t tr = (int32_t)(p[4] * dev_randnum<R>(st));
        R a = (ANG + (2.0 * M_PI) * tr) * p[2];
        R c = m_floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        R sa, ca;
        a = a * p[5] + c * p[0];
        m_sincos(a, &sa, &ca);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 78: if constexpr (vs_has(VS, 78)) { // wedge_sph
        R r = 1.0 / (N2 + EPS);
        R a = ANG + p[0] * r;
        R c = m_floor((p[1] * a + M_PI) * (M_1_PI * 0.5));
        R sa, ca;
        a = a * p[4] + c * p[2];
        m_sincos(a, &sa, &ca);
        R k = r + p[3];
        ox = ca * k;
        oy = sa * k;
    } break;
    case 79: if constexpr (vs_has(VS, 79)) { // whorl
        R r = N2;
        R a = ANG;
        a += p[r >= p[2] ? 1 : 0] / (p[2] - r);
        R sa, ca;
        m_sincos(a, &sa, &ca);
        ox = ca * r;
        oy = sa * r;
    } break;
    case 80: if constexpr (vs_has(VS, 80)) { // supershape
        R theta = p[0] * ANG + M_PI_4;
        R sn, cs;
        m_sincos(theta, &sn, &cs);
        R t1 = m_pow(m_fabs((double)cs), (double)p[2]);
        R t2 = m_pow(m_fabs((double)sn), (double)p[3]);
        R tr = N2;
        R r = (p[4] * dev_randnum<R>(st) + (1.0 - p[4]) * tr) - p[5];
        r *= m_pow((double)(t1 + t2), (double)p[1]) / tr;
        ox = x * r;
        oy = y * r;
    } break;
    case 81: if constexpr (vs_has(VS, 81)) { // flower
        R theta = ANG;
        R r = (dev_randnum<R>(st) - p[1]) * m_cos((double)(p[0] * theta));
        r /= N2 + EPS;
        ox = x * r;
        oy = y * r;
    } break;
    case 82: if constexpr (vs_has(VS, 82)) { // conic
        R tr = N2;
        R ct = x / (tr + EPS);
        R r = (dev_randnum<R>(st) - p[1]) * p[0] / (tr + tr * p[0] * ct);
        ox = x * r;
        oy = y * r;
    } break;
    case 83: if constexpr (vs_has(VS, 83)) { // parabola
        R sr, cr;
        m_sincos(N2, &sr, &cr);
        R a = p[0] * sr * sr * dev_randnum<R>(st);
        R b = p[1] * cr * dev_randnum<R>(st);
        ox = a;
        oy = b;
    } break;
    case 84: if constexpr (vs_has(VS, 84)) { // bipolar
        R x2y2 = n2sq;
        R t = x2y2 + 1.0;
        R x2 = 2.0 * x;
        R yy = 0.5 * m_atan2(2.0 * y, x2y2 - 1.0) + p[0];
        yy -= M_PI * m_floor(yy * M_1_PI + 0.5);
        ox = m_log((t + x2) / (t - x2));
        oy = yy;
    } break;
    case 85: if constexpr (vs_has(VS, 85)) { // boarders
        R rx = m_rint(x), ry = m_rint(y);
        R ox0 = x - rx, oy0 = y - ry;
        if (dev_randnum<R>(st) >= p[0]) {
            ox = ox0 * 0.5 + rx;
            oy = oy0 * 0.5 + ry;
        } else {
            R mag = 1.0 - p[0];
            if (m_fabs((double)ox0) >= m_fabs((double)oy0)) {
                R sg = m_copysign((double)mag, (double)ox0);
                ox = ox0 * 0.5 + rx + sg;
                oy = oy0 * 0.5 + ry + sg * oy0 / ox0;
            } else {
                R sg = m_copysign((double)mag, (double)oy0);
                ox = ox0 * 0.5 + rx + sg * ox0 / oy0;
                oy = oy0 * 0.5 + ry + sg;
            }
        }
    } break;
    case 86: if constexpr (vs_has(VS, 86)) { // butterfly
        R y2 = 2.0 * y;
        R r = m_sqrt(m_fabs((double)(x * y)) / (x * x + y2 * y2 + EPS));
        ox = x * r;
        oy = y2 * r;
    } break;
    case 87: if constexpr (vs_has(VS, 87)) { // cell
        R cx = m_floor((double)(x * p[1]));
        R cy = m_floor((double)(y * p[1]));
        R dx = x - cx * p[0];
        R dy = y - cy * p[0];
        R xs = m_copysign(2.0, (double)cx);
        R ys = m_copysign(2.0, (double)cy);
        R x2 = cx * xs;
        R y2 = cy * ys;
        x2 -= (R)(cx < 0);
        y2 -= (R)(cy < 0);
        ox = dx + x2 * p[0];
        oy = -dy - y2 * p[0];
    } break;
    case 88: if constexpr (vs_has(VS, 88)) { // cpow
        R a = ANG;
        R lnr = 0.5 * m_log((double)n2sq);
        R ang = p[1] * a + p[2] * lnr + p[0] * m_floor((double)(p[3] * dev_randnum<R>(st)));
        R sa, ca;
        m_sincos(ang, &sa, &ca);
        R k = m_exp((double)(p[1] * lnr - p[2] * a));
        ox = ca * k;
        oy = sa * k;
    } break;
    case 89: if constexpr (vs_has(VS, 89)) { // curve
        R vx = p[0] * m_exp((double)(-y * y * p[2]));
        R vy = p[1] * m_exp((double)(-x * x * p[3]));
        ox = x + vx;
        oy = y + vy;
    } break;
    case 90: if constexpr (vs_has(VS, 90)) { // edisc
        R tmp = n2sq + 1.0;
        R tmp2 = 2.0 * x;
        R xmax = 0.5 * (m_sqrt((double)(tmp + tmp2)) + m_sqrt((double)(tmp - tmp2)));
        R a1 = m_log(xmax + m_sqrt(xmax - 1.0));
        R a2 = -m_acos((double)(x / xmax));
        R s1, c1;
        m_sincos(a1, &s1, &c1);
        R s2 = m_sinh((double)a2), c2 = m_cosh((double)a2);
        s1 *= m_copysign(1.0, -(double)y);
        ox = c2 * c1;
        oy = s2 * s1;
    } break;
    case 91: if constexpr (vs_has(VS, 91)) { // elliptic
        R tmp = n2sq + 1.0;
        R x2 = 2.0 * x;
        R xmax = 0.5 * (m_sqrt((double)(tmp + x2)) + m_sqrt((double)(tmp - x2)));
        R a = x / xmax;
        R b = 1.0 - a * a;
        R ssx = xmax - 1.0;
        b = b < 0.0 ? (R)0.0 : (R)m_sqrt((double)b);
        ssx = ssx < 0.0 ? (R)0.0 : (R)m_sqrt((double)ssx);
        ox = m_atan2((double)a, (double)b);
        oy = m_copysign(1.0, (double)y) * m_log((double)(xmax + ssx));
    } break;
    case 92: if constexpr (vs_has(VS, 92)) { // escher
        R a = ANG;
        R lnr = 0.5 * m_log((double)n2sq);
        R n = p[0] * a + p[1] * lnr;
        R sn, cn;
        m_sincos(n, &sn, &cn);
        R k = m_exp((double)(p[0] * lnr - p[1] * a));
        ox = cn * k;
        oy = sn * k;
    } break;
    case 93: if constexpr (vs_has(VS, 93)) { // foci
        R expx = 0.5 * m_exp((double)x);
        R expnx = 0.25 / expx;
        R sn, cn;
        m_sincos(y, &sn, &cn);
        R tmp = 1.0 / (expx + expnx - cn);
        R a = expx - expnx;
        ox = a * tmp;
        oy = sn * tmp;
    } break;
    case 94: if constexpr (vs_has(VS, 94)) { // lazysusan
        R lx = x - p[0];
        R ly = y + p[1];
        R r = m_hypot((double)lx, (double)ly);
        if (r < p[5]) {
            R a = m_atan2((double)ly, (double)lx) + p[2] + p[3] * (p[5] - r);
            R sa, ca;
            m_sincos(a, &sa, &ca);
            ox = r * ca + p[0];
            oy = r * sa - p[1];
        } else {
            r = 1.0 + p[4] / (r + EPS);
            ox = r * lx + p[0];
            oy = r * ly - p[1];
        }
    } break;
    case 95: if constexpr (vs_has(VS, 95)) { // loonie
        R r2 = n2sq;
        R w2 = p[1];
        R r = p[0];
        if (r2 < w2)
            r *= m_sqrt(w2 / (r2 + EPS) - 1.0);
        ox = x * r;
        oy = y * r;
    } break;
    case 96: if constexpr (vs_has(VS, 96)) { // oscope
        R damp = m_exp(-m_fabs((double)x) * p[2]);
        R t = p[1] * damp * m_cos((double)(p[0] * x)) + p[3];
        R yy = m_copysign(1.0, m_fabs((double)y) - t) * y;
        ox = x;
        oy = yy;
    } break;
    case 97: if constexpr (vs_has(VS, 97)) { // popcorn
        R dx = p[0] * m_sin(m_tan((double)(y * p[2])));
        R dy = p[1] * m_sin(m_tan((double)(x * p[2])));
        ox = x + dx;
        oy = y + dy;
    } break;
    default:
        break;
    }
#undef N2
#undef ANG
}

// N-D variations and the VariationFrom2D embedding (variations.hpp:63-107,
// 160-505, 2376-2386).  out must start at zero.
template <typename R, int D, int VS, bool NATIVE>
__device__ __forceinline__ void var_eval(int type, const R *E, const R *t, R *out, Stream &st)
{
    const R EPS = Consts<R>::eps;
    if (type >= 20) {
        if constexpr (D >= 2) {
            if constexpr (D == 2) {
                R ox = 0, oy = 0;
                calc2d<R, VS, NATIVE>(type, E + 1, t[0], t[1], ox, oy, st);
                out[0] = ox;
                out[1] = oy;
            } else {
                int ax = (int)E[1], ay = (int)E[2];
                R xx = ax == 0 ? t[0] : (ax == 1 ? t[1] : t[2]);
                R yy = ay == 0 ? t[0] : (ay == 1 ? t[1] : t[2]);
                R ox = 0, oy = 0;
                calc2d<R, VS, NATIVE>(type, E + 3, xx, yy, ox, oy, st);
#pragma unroll
                for (int i = 0; i < D; ++i)
                    out[i] = (i == ax) ? ox : ((i == ay) ? oy : (R)0);
            }
        }
        return;
    }
    const R *p = E + 1;
    switch (type) {
    case 0: // linear
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i];
        break;
    case 1: if constexpr (vs_has(VS, 1)) { // sinusoidal
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = m_sin(t[i]);
    } break;
    case 2: if constexpr (vs_has(VS, 2)) { // spherical
        R r = recip_ref(dev_norm2sq(t, D) + EPS);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 3: if constexpr (vs_has(VS, 3)) { // bent
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R c = t[i];
            if (c < 0.0)
                c *= p[i];
            else
                c *= p[3 + i];
            out[i] = c;
        }
    } break;
    case 4: if constexpr (vs_has(VS, 4)) { // rectangles
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R q = p[i], x = t[i];
            out[i] = (q == 0.0) ? x : (R)((2.0 * m_floor((double)(x / q)) + 1.0) * q - x);
        }
    } break;
    case 5: if constexpr (vs_has(VS, 5)) { // fisheye
        R n = D == 1 ? (R)m_fabs(t[0]) : (R)m_sqrt(dev_norm2sq(t, D));
        R r = 1.0 / (n + p[0]);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 6: if constexpr (vs_has(VS, 6)) { // bubble
        R r = 1.0 / (dev_norm2sq(t, D) + p[0]);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 7: if constexpr (vs_has(VS, 7)) { // noise
        R r = dev_randnum<R>(st);
        R dir[D];
        dev_randdir<R, D>(st, dir);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R m = t[i] * dir[i];
            out[i] = m * r;
        }
    } break;
    case 8: if constexpr (vs_has(VS, 8)) { // blur
        R r = dev_randnum<R>(st);
        R dir[D];
        dev_randdir<R, D>(st, dir);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = dir[i] * r;
    } break;
    case 9:  // gaussian_blur
    case 13: if constexpr (vs_has(VS, 9)) { // pre_blur
        R g = dev_randgauss<R>(st);
        R dir[D];
        dev_randdir<R, D>(st, dir);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = dir[i] * g;
    } break;
    case 10: if constexpr (vs_has(VS, 10)) { // square_noise
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = dev_randnum<R>(st) - 0.5;
    } break;
    case 11: if constexpr (vs_has(VS, 11)) { // separation
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double sg = m_copysign(1.0, (double)t[i]);
            out[i] = sg * (m_sqrt((double)(t[i] * t[i] + p[i])) - sg * p[3 + i]);
        }
    } break;
    case 12: if constexpr (vs_has(VS, 12)) { // splits
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R q = m_copysign(1.0, (double)t[i]) * p[i];
            out[i] = t[i] + q;
        }
    } break;
    case 14: if constexpr (vs_has(VS, 14)) { // modulus
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] - p[i] * m_floor(t[i] * p[3 + i] + 0.5);
    } break;
    case 15: if constexpr (vs_has(VS, 15)) { // celln
#pragma unroll
        for (int i = 0; i < D; ++i) {
            R x = m_floor((double)(t[i] * p[3 + i]));
            R dx = t[i] - x * p[i];
            R xs = m_copysign(2.0, (double)x);
            R x2 = x * xs;
            x2 -= (R)(x < 0);
            out[i] = dx + x2 * p[i];
        }
    } break;
    case 16: if constexpr (vs_has(VS, 16)) { // spherical_p
        R ns = m_pow(m_fabs((double)t[0]), (double)p[0]);
        for (int i = 1; i < D; ++i)
            ns += m_pow(m_fabs((double)t[i]), (double)p[0]);
        R r = 1.0 / (ns + EPS);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 17: if constexpr (vs_has(VS, 17)) { // unit_sphere
        R n = D == 1 ? (R)m_fabs(t[0]) : (R)m_sqrt(dev_norm2sq(t, D));
        R r = 1.0 / (n + EPS);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 18: if constexpr (vs_has(VS, 18)) { // unit_sphere_p
        R ns = m_pow(m_fabs((double)t[0]), (double)p[0]);
        for (int i = 1; i < D; ++i)
            ns += m_pow(m_fabs((double)t[i]), (double)p[0]);
        R nrm = m_pow((double)ns, 1.0 / p[0]);
        R r = 1.0 / (nrm + EPS);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    case 19: if constexpr (vs_has(VS, 19)) { // unit_cube
        R n = m_fabs(t[0]);
        for (int i = 1; i < D; ++i) {
            R v = m_fabs(t[i]);
            n = (n < v) ? v : n;
        }
        R r = 1.0 / (n + EPS);
#pragma unroll
        for (int i = 0; i < D; ++i)
            out[i] = t[i] * r;
    } break;
    default:
        break;
    }
}

} // namespace ffr

// Flame JSON validation and compilation to the device table.
#include "flame_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ffr {

// ------------------------------------------------------------------ catalogue
// Names and parameter keys follow the reference's factory
// (variations.hpp:2387-2612); N-D ids first, then the 2-D ones.
static const std::vector<VariationInfo> &catalogue()
{
    static const std::vector<VariationInfo> cat = {
        {"linear", {}},
        {"sinusoidal", {}},
        {"spherical", {}},
        {"bent", {{"scales_neg", 'p'}, {"scales_pos", 'p'}}},
        {"rectangles", {{"params", 'p'}}},
        {"fisheye", {{"addval", 's'}}},
        {"bubble", {{"addval", 's'}}},
        {"noise", {}},
        {"blur", {}},
        {"gaussian_blur", {}},
        {"square_noise", {}},
        {"separation", {{"params", 'p'}, {"inside", 'p'}}},
        {"splits", {{"params", 'p'}}},
        {"pre_blur", {}},
        {"modulus", {{"params", 'p'}}},
        {"celln", {{"sizes", 'p'}}},
        {"spherical_p", {{"norm", 's'}}},
        {"unit_sphere", {}},
        {"unit_sphere_p", {{"norm", 's'}}},
        {"unit_cube", {}},
        {"swirl", {}},
        {"horseshoe", {}},
        {"polar", {}},
        {"polar2", {}},
        {"handkerchief", {}},
        {"heart", {}},
        {"disc", {}},
        {"disc2", {{"rotation", 's'}, {"twist", 's'}}},
        {"waves", {{"xfreq", 's'}, {"xscale", 's'}, {"yfreq", 's'}, {"yscale", 's'}}},
        {"fan", {{"x", 's'}, {"y", 's'}}},
        {"rings", {{"value", 's'}}},
        {"spiral", {}},
        {"hyperbolic", {}},
        {"diamond", {}},
        {"ex", {}},
        {"julia", {}},
        {"exponential", {}},
        {"power", {}},
        {"cosine", {}},
        {"blob", {{"low", 's'}, {"high", 's'}, {"waves", 's'}}},
        {"pdj", {{"a", 's'}, {"b", 's'}, {"c", 's'}, {"d", 's'}}},
        {"cylinder", {}},
        {"perspective", {{"distance", 's'}, {"angle", 's'}}},
        {"julian", {{"power", 's'}, {"dist", 's'}}},
        {"juliascope", {{"power", 's'}, {"dist", 's'}}},
        {"radial_blur", {{"angle", 's'}, {"flam3_weight", 's'}}},
        {"pie", {{"slices", 's'}, {"rotation", 's'}, {"thickness", 's'}}},
        {"ngon", {{"sides", 's'}, {"power", 's'}, {"corners", 's'}, {"circle", 's'}}},
        {"curl", {{"c1", 's'}, {"c2", 's'}}},
        {"arch", {{"flam3_weight", 's'}}},
        {"tangent", {}},
        {"rays", {{"flam3_weight", 's'}}},
        {"blade", {{"flam3_weight", 's'}}},
        {"secant", {{"flam3_weight", 's'}}},
        {"twintrian", {{"flam3_weight", 's'}}},
        {"cross", {}},
        {"exp", {}},
        {"log", {}},
        {"sin", {}},
        {"cos", {}},
        {"tan", {}},
        {"sec", {}},
        {"csc", {}},
        {"cot", {}},
        {"sinh", {}},
        {"cosh", {}},
        {"tanh", {}},
        {"sech", {}},
        {"csch", {}},
        {"coth", {}},
        {"auger", {{"freq", 's'}, {"flam3_weight", 's'}, {"scale", 's'}, {"sym", 's'}}},
        {"flux", {{"spread", 's'}, {"flam3_weight", 's'}}},
        {"mobius", {{"a", '2'}, {"b", '2'}, {"c", '2'}, {"d", '2'}}},
        {"scry", {{"flam3_weight", 's'}}},
        {"split", {{"xsize", 's'}, {"ysize", 's'}}},
        {"stripes", {{"space", 's'}, {"warp", 's'}}},
        {"wedge", {{"swirl", 's'}, {"count", 's'}, {"angle", 's'}, {"hole", 's'}}},
        {"wedge_julia", {{"angle", 's'}, {"count", 's'}, {"power", 's'}, {"dist", 's'}}},
        {"wedge_sph", {{"angle", 's'}, {"count", 's'}, {"swirl", 's'}, {"hole", 's'}}},
        {"whorl", {{"inside", 's'}, {"outside", 's'}, {"flam3_weight", 's'}}},
        {"supershape", {{"n1", 's'}, {"m", 's'}, {"n2", 's'}, {"n3", 's'}, {"rnd", 's'}, {"holes", 's'}}},
        {"flower", {{"petals", 's'}, {"holes", 's'}}},
        {"conic", {{"eccen", 's'}, {"holes", 's'}}},
        {"parabola", {{"height", 's'}, {"width", 's'}}},
        {"bipolar", {{"shift", 's'}}},
        {"boarders", {{"prob", 's'}}},
        {"butterfly", {}},
        {"cell", {{"size", 's'}}},
        {"cpow", {{"r", 's'}, {"i", 's'}, {"power", 's'}}},
        {"curve", {{"xamp", 's'}, {"yamp", 's'}, {"xlen", 's'}, {"ylen", 's'}}},
        {"edisc", {}},
        {"elliptic", {}},
        {"escher", {{"beta", 's'}}},
        {"foci", {}},
        {"lazysusan", {{"x", 's'}, {"y", 's'}, {"spin", 's'}, {"twist", 's'}, {"space", 's'}, {"flam3_weight", 's'}}},
        {"loonie", {{"flam3_weight", 's'}}},
        {"oscope", {{"frequency", 's'}, {"amplitude", 's'}, {"damping", 's'}, {"separation", 's'}}},
        {"popcorn", {{"x", 's'}, {"y", 's'}, {"c", 's'}}},
    };
    return cat;
}

const VariationInfo &variation_info(int type) { return catalogue().at((size_t)type); }

int variation_id(const std::string &name)
{
    const auto &c = catalogue();
    for (size_t i = 0; i < c.size(); ++i)
        if (name == c[i].name)
            return (int)i;
    return -1;
}

int variation_draws(int type, int dims)
{
    int dir = dims == 3 ? 2 : 1; // randDirection<dims> words (flame_rng.hpp:171-206)
    switch (type) {
    case 7:  /* noise */
    case 8:  /* blur */ return 1 + dir;
    case 9:  /* gaussian_blur */
    case 13: /* pre_blur */ return 2 + dir;
    case 10: /* square_noise */ return dims;
    case 35: /* julia */
    case 43: /* julian */
    case 49: /* arch */
    case 51: /* rays */
    case 52: /* blade */
    case 54: /* twintrian */
    case 77: /* wedge_julia */
    case 80: /* supershape */
    case 81: /* flower */
    case 82: /* conic */
    case 85: /* boarders */
    case 88: /* cpow */ return 1;
    case 44: /* juliascope */
    case 45: /* radial_blur */
    case 83: /* parabola */ return 2;
    case 46: /* pie */ return 3;
    default: return 0;
    }
}

uint64_t FlameSpec::cells() const
{
    uint64_t n = 1;
    for (int i = 0; i < dims; ++i)
        n *= size[i];
    return n;
}

// ------------------------------------------------------------------ parsing

static double to_num(double v, bool fp32) { return fp32 ? (double)(float)v : v; }

static std::vector<double> parse_point(const Json &j, int n)
{
    // Point(Json&) (point.hpp:29-47)
    if (!j.isArray())
        throw JsonError("Point(Json&): not an array");
    const auto &a = j.arrayValue();
    if (a.size() != (size_t)n)
        throw JsonError("Point(Json&): incorrect array size: " + std::to_string(a.size()));
    std::vector<double> out;
    for (const auto &e : a) {
        if (e.isInt())
            out.push_back((double)e.intValue());
        else if (e.isFloat())
            out.push_back(e.floatValue());
        else
            throw JsonError("Point(Json&): entry is not a number");
    }
    return out;
}

static void parse_affine(const Json &j, int n, double *A, double *b)
{
    // Affine(Json&) (affine.hpp:45-92)
    if (!j.isObject())
        throw JsonError("Affine(Json&): not an object");
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
            A[r * n + c] = (r == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i)
        b[i] = 0.0;
    if (j.has("A")) {
        const Json &ja = j["A"];
        if (!ja.isArray())
            throw JsonError("Affine(Json&): A is not an array");
        if (ja.size() != (size_t)n)
            throw JsonError("Affine(Json&): A is wrong size");
        for (int i = 0; i < n; ++i) {
            try {
                auto row = parse_point(ja[(size_t)i], n);
                for (int c = 0; c < n; ++c)
                    A[i * n + c] = row[(size_t)c];
            } catch (const std::exception &e) {
                throw JsonError("Affine(Json&): error parsing A[" + std::to_string(i) +
                                "]: " + e.what());
            }
        }
    }
    if (j.has("b")) {
        try {
            auto bb = parse_point(j["b"], n);
            for (int i = 0; i < n; ++i)
                b[i] = bb[(size_t)i];
        } catch (const std::exception &e) {
            throw JsonError(std::string("Affine(Json&): error parsing b: ") + e.what());
        }
    }
}

static VarSpec parse_variation(const Json &j, int dims)
{
    // Variation<dims>::parseVariation (variations.hpp:2614-2626)
    std::string name = j["name"].stringValue();
    int id = variation_id(name);
    if (id < 0 || (dims < 2 && id >= kNumNdVariations))
        throw JsonError("unknown variation: " + name);
    VarSpec v;
    v.type = id;
    v.weight = j["weight"].floatValue();
    if (id >= kNumNdVariations && dims > 2) {
        // VariationFrom2D axis selection (variations.hpp:70-89)
        int64_t ax = j["axis_x"].intValue();
        int64_t ay = j["axis_y"].intValue();
        if (ax < 0 || ax >= dims)
            throw JsonError("axis_x index out of range");
        if (ay < 0 || ay >= dims)
            throw JsonError("axis_y index out of range");
        if (ax == ay)
            throw JsonError("axes are not distinct");
        v.axis_x = (int)ax;
        v.axis_y = (int)ay;
    }
    for (const auto &kv : variation_info(id).keys) {
        const Json &val = j[kv.first];
        if (kv.second == 's') {
            v.params.push_back(val.floatValue());
        } else {
            auto pt = parse_point(val, kv.second == 'p' ? dims : 2);
            v.params.insert(v.params.end(), pt.begin(), pt.end());
        }
    }
    if (name == "boarders" && (v.params[0] < 0.0 || v.params[0] > 1.0))
        throw JsonError("boarders probability out of range");
    if ((name == "spherical_p" || name == "unit_sphere_p") && v.params[0] <= 0.0)
        throw JsonError("norm <= 0");
    return v;
}

static XformSpec parse_xform(const Json &j, int id, bool is_final, int dims,
                             int color_dims, bool fp32)
{
    // XForm(Json&, ...) (xform.hpp:71-172)
    XformSpec x;
    x.id = id;
    if (!is_final) {
        try {
            x.weight = j["weight"].floatValue();
        } catch (const std::exception &e) {
            throw JsonError(std::string("XForm(): cannot parse weight: ") + e.what());
        }
    } else {
        x.weight = 1.0;
    }
    if (to_num(x.weight, fp32) < 0.0)
        throw JsonError("XForm(): weight is negative");
    if (j.has("pre_affine")) {
        try {
            parse_affine(j["pre_affine"], dims, x.preA, x.preb);
        } catch (const std::exception &e) {
            throw JsonError(std::string("XForm(): cannot parse pre_affine") + e.what());
        }
        x.has_pre = true;
    } else {
        parse_affine(Json::parse("{}"), dims, x.preA, x.preb);
    }
    if (j.has("post_affine")) {
        try {
            parse_affine(j["post_affine"], dims, x.postA, x.postb);
        } catch (const std::exception &e) {
            throw JsonError(std::string("XForm(): cannot parse post_affine") + e.what());
        }
        x.has_post = true;
    } else {
        parse_affine(Json::parse("{}"), dims, x.postA, x.postb);
    }
    try {
        for (const auto &vj : j["variations"].arrayValue())
            x.vars.push_back(parse_variation(vj, dims));
    } catch (const std::exception &e) {
        throw JsonError(std::string("XForm(): cannot parse variations: ") + e.what());
    }
    if (color_dims && j.has("color")) {
        const auto &ca = j["color"].arrayValue(); // outside the try, as the reference
        try {
            if (ca.size() != (size_t)color_dims)
                throw JsonError("color length incorrect");
            for (const auto &c : ca) {
                double v = c.floatValue();
                double vn = to_num(v, fp32);
                if (vn < 0.0 || vn > 1.0)
                    throw JsonError("color coordinate out of range");
                x.color.push_back(v);
            }
        } catch (const std::exception &e) {
            throw JsonError(std::string("XForm(): cannot parse color: ") + e.what());
        }
        x.has_color = true;
    }
    if (color_dims && j.has("color_speed")) {
        try {
            x.speed = j["color_speed"].floatValue();
            double sn = to_num(x.speed, fp32);
            if (sn < 0.0 || sn > 1.0)
                throw JsonError("color speed out of range");
        } catch (const std::exception &e) {
            throw JsonError(std::string("XForm(): color speed issue: ") + e.what());
        }
        x.has_speed = true;
    }
    return x;
}

int flame_dims(const Json &doc) { return (int)doc["dimensions"].intValue(); }

static uint64_t to_size_t(double v)
{
    if (!(v > -1.0) || v >= 18446744073709551616.0)
        return 0;
    return v >= 0 ? (uint64_t)v : 0;
}

FlameSpec parse_flame(const Json &doc, bool fp32)
{
    FlameSpec f;
    f.dims = flame_dims(doc);
    if (f.dims < 1 || f.dims > 3)
        throw JsonError("ERROR: " + std::to_string(f.dims) + "D not supported");
    const int dims = f.dims;
    const double M = fp32 ? 1e5 : 1e10; // max_rect (constants.hpp:32-35)
    try {
        const auto &sizej = doc["size"].arrayValue();
        const auto &boundsj = doc["bounds"].arrayValue();
        if (sizej.size() != (size_t)dims)
            throw JsonError("incorrect size length");
        if (boundsj.size() != (size_t)dims)
            throw JsonError("incorrect bounds length");
        for (int i = 0; i < dims; ++i) {
            double sv;
            try {
                sv = sizej[(size_t)i].floatValue();
            } catch (const std::exception &e) {
                throw JsonError("error parsing size[" + std::to_string(i) + "]: " + e.what());
            }
            f.size[i] = to_size_t(sv);
            if (f.size[i] == 0 || f.size[i] > kMaxDim)
                throw JsonError("size[" + std::to_string(i) + "] out of range");
            const auto &bj = boundsj[(size_t)i].arrayValue();
            if (bj.size() != 2)
                throw JsonError("bounds[" + std::to_string(i) + "] wrong format");
            double lo, hi;
            try {
                lo = bj[0].floatValue();
                hi = bj[1].floatValue();
            } catch (const std::exception &e) {
                throw JsonError("error parsing bounds[" + std::to_string(i) + "]: " + e.what());
            }
            f.lo[i] = lo;
            f.hi[i] = hi;
            double ln = to_num(lo, fp32), hn = to_num(hi, fp32);
            if (ln < -M || ln > M || hn < -M || hn > M)
                throw JsonError("bounds[" + std::to_string(i) + "] out of range");
            if (ln >= hn)
                throw JsonError("bounds[" + std::to_string(i) + "] low >= high");
        }
    } catch (const std::exception &e) {
        throw JsonError(std::string("Flame(): ") + e.what());
    }
    f.has_final = doc.has("final_xform");
    const std::vector<Json> *xfs = nullptr;
    try {
        xfs = &doc["xforms"].arrayValue();
    } catch (const std::exception &e) {
        throw JsonError(std::string("Flame(): cannot parse xforms: ") + e.what());
    }
    try {
        if (doc.has("color_dimensions")) {
            int64_t cd = doc["color_dimensions"].intValue();
            if (cd < 0 || cd > kMaxColorDims)
                throw JsonError("too many color dimensions");
            f.color_dims = (int)cd;
        }
        if (doc.has("color_speed"))
            f.color_speed = doc["color_speed"].floatValue();
        double cs = to_num(f.color_speed, fp32);
        if (cs < 0.0 || cs > 1.0)
            throw JsonError("color speed out of range");
    } catch (const std::exception &e) {
        throw JsonError(std::string("Flame(): ") + e.what());
    }
    int id = 0;
    for (const auto &xj : *xfs) {
        try {
            f.xforms.push_back(parse_xform(xj, id, false, dims, f.color_dims, fp32));
        } catch (const std::exception &e) {
            throw JsonError("Flame(): error parsing xforms[" + std::to_string(id) +
                            "]: " + e.what());
        }
        ++id;
    }
    if (f.xforms.empty())
        throw JsonError("Flame(): no xforms");
    if (f.has_final) {
        try {
            f.final_xf = parse_xform(doc["final_xform"], -1, true, dims, f.color_dims, fp32);
        } catch (const std::exception &e) {
            throw JsonError(std::string("Flame(): error parsing final xform: ") + e.what());
        }
    }
    bool any = false;
    for (const auto &x : f.xforms)
        if (to_num(x.weight, fp32) != 0.0)
            any = true;
    if (!any)
        throw JsonError("Flame(): no xforms remaining after optimization");
    uint64_t cells = 1;
    for (int i = 0; i < dims; ++i) {
        cells *= f.size[i];
        if (cells >= kMaxCells)
            throw JsonError("BufferRenderer(): histogram too big");
    }
    return f;
}

// ------------------------------------------------------------------ compile

XformLayout xform_layout(int dims, int color_dims)
{
    XformLayout L;
    int o = 0;
    L.preA = o; o += dims * dims;
    L.preb = o; o += dims;
    L.postA = o; o += dims * dims;
    L.postb = o; o += dims;
    L.speed = o; o += 1;
    L.has_color = o; o += 1;
    L.has_pre = o; o += 1;
    L.has_post = o; o += 1;
    L.color = o; o += color_dims;
    L.stride = (o + 1) & ~1;
    return L;
}

// number of precomputed values a variation keeps (the member fields of its
// reference class after construction)
static int var_nparams(int type, int dims)
{
    switch (type) {
    case 3: case 11: return 6;          // bent, separation
    case 4: case 12: return 3;          // rectangles, splits
    case 14: case 15: return 6;         // modulus, celln
    case 5: case 6: case 16: case 18: return 1;
    case 27: return 3;                  // disc2
    case 28: return 4;                  // waves
    case 29: return 2;                  // fan
    case 30: return 1;                  // rings
    case 39: return 3;                  // blob
    case 40: return 4;                  // pdj
    case 42: return 3;                  // perspective
    case 43: case 44: return 3;         // julian, juliascope
    case 45: return 3;                  // radial_blur
    case 46: return 4;                  // pie
    case 47: return 5;                  // ngon
    case 48: return 2;                  // curl
    case 49: case 51: case 52: case 53: case 54: case 73: return 1;
    case 95: return 2;                  // loonie
    case 70: return 4;                  // auger
    case 71: return 2;                  // flux
    case 72: return 8;                  // mobius
    case 74: case 75: return 2;         // split, stripes
    case 76: case 78: return 5;         // wedge, wedge_sph
    case 77: return 6;                  // wedge_julia
    case 79: return 3;                  // whorl
    case 80: return 6;                  // supershape
    case 81: case 82: case 83: return 2; // flower, conic, parabola
    case 84: case 85: return 1;         // bipolar, boarders
    case 87: return 2;                  // cell
    case 88: return 4;                  // cpow
    case 89: return 4;                  // curve
    case 92: return 2;                  // escher
    case 94: return 6;                  // lazysusan
    case 96: return 4;                  // oscope
    case 97: return 3;                  // popcorn
    default: (void)dims; return 0;
    }
}

// the reference constructors' arithmetic, in num_t (variations.hpp)
template <typename R> static void var_setup(const VarSpec &v, int dims, R *p)
{
    const std::vector<double> &q = v.params;
    const double kPi = M_PI, kPi2 = M_PI_2, k1Pi = M_1_PI;
    const R eps = sizeof(R) == 4 ? (R)1e-10f : (R)1e-20;
    switch (v.type) {
    case 3: case 11:
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[(size_t)i];
            p[3 + i] = (R)q[(size_t)(dims + i)];
        }
        if (v.type == 11)
            for (int i = 0; i < dims; ++i)
                p[i] *= p[i];
        break;
    case 4: case 12:
        for (int i = 0; i < dims; ++i)
            p[i] = (R)q[(size_t)i];
        break;
    case 5: case 6: case 16: case 18:
        p[0] = (R)q[0];
        break;
    case 14:
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[(size_t)i];
            p[i] *= (R)2.0;
            p[3 + i] = (R)(1.0 / p[i]);
        }
        break;
    case 15:
        for (int i = 0; i < dims; ++i) {
            p[i] = (R)q[(size_t)i];
            p[3 + i] = (R)(1.0 / p[i]);
        }
        break;
    case 27: {
        R rot = (R)q[0], twist = (R)q[1];
        p[0] = (R)(rot * kPi);
        R sinadd, cosadd;
        if (sizeof(R) == 4) {
            float s, c;
            ::sincosf((float)twist, &s, &c);
            sinadd = (R)s;
            cosadd = (R)c;
        } else {
            double s, c;
            ::sincos((double)twist, &s, &c);
            sinadd = (R)s;
            cosadd = (R)c;
        }
        cosadd = (R)(cosadd - 1.0);
        R k = (R)(1.0 + twist);
        if (twist > 2.0 * kPi)
            k = (R)(k - 2.0 * kPi);
        if (twist < -2.0 * kPi)
            k = (R)(k + 2.0 * kPi);
        cosadd *= k;
        sinadd *= k;
        p[1] = cosadd;
        p[2] = sinadd;
        break;
    }
    case 28: case 40:
        for (int i = 0; i < 4; ++i)
            p[i] = (R)q[(size_t)i];
        break;
    case 29: {
        R x = (R)q[0], y = (R)q[1];
        p[0] = (R)(kPi * (double)(x * x + eps));
        p[1] = y;
        break;
    }
    case 30: {
        R val = (R)q[0];
        p[0] = val * val + eps;
        break;
    }
    case 39: {
        R low = (R)q[0], high = (R)q[1];
        p[0] = (R)((high + low) / 2.0);
        p[1] = (R)((high - low) / 2.0);
        p[2] = (R)q[2];
        break;
    }
    case 42: {
        p[0] = (R)q[0];
        R angle = (R)q[1];
        p[1] = (R)::sin((double)angle);
        p[2] = (R)(p[0] * ::cos((double)angle));
        break;
    }
    case 43: case 44: {
        R power = (R)q[0], dist = (R)q[1];
        p[0] = (R)::fabs((double)power);
        p[1] = (R)(1.0 / power);
        p[2] = (R)(dist / (2.0 * power));
        break;
    }
    case 45: {
        R angle = (R)q[0];
        if (sizeof(R) == 4) {
            float s, c;
            ::sincosf((float)(angle * kPi2), &s, &c);
            p[0] = (R)s;
            p[1] = (R)c;
        } else {
            double s, c;
            ::sincos((double)(angle * kPi2), &s, &c);
            p[0] = (R)s;
            p[1] = (R)c;
        }
        p[2] = (R)q[1];
        break;
    }
    case 46:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        p[3] = (R)((2.0 * kPi) / p[0]);
        break;
    case 47: {
        R sides = (R)q[0];
        p[0] = (R)(q[1] / 2.0);
        p[1] = (R)((2.0 * kPi) / sides);
        p[2] = (R)q[2];
        p[3] = (R)q[3];
        p[4] = (R)(sides / (2.0 * kPi));
        break;
    }
    case 48:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        break;
    case 49: case 51: case 52: case 53: case 54: case 73:
        p[0] = (R)q[0];
        break;
    case 95:
        p[0] = (R)q[0];
        p[1] = p[0] * p[0];
        break;
    case 70:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)(q[2] / 2.0);
        p[3] = (R)q[3];
        break;
    case 71:
        p[0] = (R)(2.0 + q[0]);
        p[1] = (R)q[1];
        break;
    case 72:
        for (int i = 0; i < 8; ++i)
            p[i] = (R)q[(size_t)i];
        break;
    case 74:
        p[0] = (R)(q[0] * kPi);
        p[1] = (R)(q[1] * kPi);
        break;
    case 75:
        p[0] = (R)(1.0 - q[0]);
        p[1] = (R)q[1];
        break;
    case 76: case 78: {
        R swirl, count, angle, hole;
        if (v.type == 76) {
            swirl = (R)q[0]; count = (R)q[1]; angle = (R)q[2]; hole = (R)q[3];
        } else {
            angle = (R)q[0]; count = (R)q[1]; swirl = (R)q[2]; hole = (R)q[3];
        }
        p[0] = swirl;
        p[1] = count;
        p[2] = angle;
        p[3] = hole;
        p[4] = (R)(1.0 - angle * count * (k1Pi * 0.5));
        break;
    }
    case 77: {
        R angle = (R)q[0], count = (R)q[1], power = (R)q[2], dist = (R)q[3];
        p[0] = angle;
        p[1] = count;
        p[2] = (R)(1.0 / power);
        p[3] = (R)(dist / (2.0 * power));
        p[4] = (R)::fabs((double)power);
        p[5] = (R)(1.0 - angle * count * (k1Pi * 0.5));
        break;
    }
    case 79:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        break;
    case 80: {
        R n1 = (R)q[0];
        p[0] = (R)(q[1] / 4.0);
        p[1] = (R)(-1.0 / n1);
        for (int i = 2; i < 6; ++i)
            p[i] = (R)q[(size_t)i];
        break;
    }
    case 81: case 82: case 83:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        break;
    case 84:
        p[0] = (R)(-kPi2 * q[0]);
        break;
    case 85:
        p[0] = (R)q[0];
        break;
    case 87:
        p[0] = (R)q[0];
        p[1] = (R)(1.0 / p[0]);
        break;
    case 88: {
        R rr = (R)q[0], ii = (R)q[1], power = (R)q[2];
        p[0] = (R)(2.0 * kPi / power);
        p[1] = (R)(rr / power);
        p[2] = (R)(ii / power);
        p[3] = power;
        break;
    }
    case 89: {
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        R xlen = (R)q[2], ylen = (R)q[3];
        xlen *= xlen;
        ylen *= ylen;
        p[2] = (R)(1.0 / std::max(eps, xlen));
        p[3] = (R)(1.0 / std::max(eps, ylen));
        break;
    }
    case 92: {
        R beta = (R)q[0];
        R seb, ceb;
        if (sizeof(R) == 4) {
            float s, c;
            ::sincosf((float)beta, &s, &c);
            seb = (R)s;
            ceb = (R)c;
        } else {
            double s, c;
            ::sincos((double)beta, &s, &c);
            seb = (R)s;
            ceb = (R)c;
        }
        p[0] = (R)(0.5 * (1.0 + ceb));
        p[1] = (R)(0.5 * seb);
        break;
    }
    case 94:
        for (int i = 0; i < 6; ++i)
            p[i] = (R)q[(size_t)i];
        break;
    case 96: {
        R freq = (R)q[0];
        p[0] = (R)(2.0 * kPi * freq);
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        p[3] = (R)q[3];
        break;
    }
    case 97:
        p[0] = (R)q[0];
        p[1] = (R)q[1];
        p[2] = (R)q[2];
        break;
    default:
        break;
    }
}

// shortest common supersequence of two type sequences (LCS dynamic program)
static std::vector<int> scs_merge(const std::vector<int> &a, const std::vector<int> &b)
{
    size_t n = a.size(), m = b.size();
    std::vector<std::vector<int>> L(n + 1, std::vector<int>(m + 1, 0));
    for (size_t i = n; i-- > 0;)
        for (size_t j = m; j-- > 0;)
            L[i][j] = (a[i] == b[j]) ? 1 + L[i + 1][j + 1] : std::max(L[i + 1][j], L[i][j + 1]);
    std::vector<int> out;
    size_t i = 0, j = 0;
    while (i < n && j < m) {
        if (a[i] == b[j]) {
            out.push_back(a[i]);
            ++i;
            ++j;
        } else if (L[i + 1][j] >= L[i][j + 1]) {
            out.push_back(a[i++]);
        } else {
            out.push_back(b[j++]);
        }
    }
    while (i < n)
        out.push_back(a[i++]);
    while (j < m)
        out.push_back(b[j++]);
    return out;
}

template <typename R> DeviceFlame<R> compile_flame(const FlameSpec &spec)
{
    DeviceFlame<R> d;
    const int dims = spec.dims, r = spec.color_dims;
    d.dims = dims;
    d.r = r;
    d.n_ids = (int)spec.xforms.size();
    d.has_final = spec.has_final ? 1 : 0;
    const R eps = sizeof(R) == 4 ? (R)1e-10f : (R)1e-20;
    (void)eps;

    // Flame::_optimize: drop weight-0 xforms, std::sort by decreasing weight
    struct Item {
        R w;
        int idx;
    };
    std::vector<Item> items;
    for (size_t i = 0; i < spec.xforms.size(); ++i) {
        R w = (R)spec.xforms[i].weight;
        if (w == (R)0.0)
            continue;
        items.push_back({w, (int)i});
    }
    if (items.empty())
        throw JsonError("Flame(): no xforms remaining after optimization");
    std::sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.w > b.w; });
    d.nxf = (int)items.size();
    for (const auto &it : items)
        d.order.push_back(it.idx);
    // Flame::_setupCumulativeWeights
    R normdiv = 0.0;
    for (const auto &it : items)
        normdiv += it.w;
    R wsum = 0.0;
    for (const auto &it : items) {
        wsum += it.w / normdiv;
        d.cw.push_back(wsum);
    }
    d.cw.back() = (R)1.0;
    // BufferRenderer::_init
    const R down = sizeof(R) == 4 ? (R)(1.0f - 1.0f / (float)(1 << 23))
                                  : (R)(1.0 - (double)(1.0f / (float)(1LL << 52)));
    d.cells = 1;
    for (int i = 0; i < dims; ++i) {
        d.lo[i] = (R)spec.lo[i];
        d.hi[i] = (R)spec.hi[i];
        d.mult_d[i] = (R)(spec.size[i]) / (d.hi[i] - d.lo[i]);
        d.mult_d[i] *= down;
        d.mult_i[i] = d.cells;
        d.cells *= spec.size[i];
    }

    // variation lists after XForm::_optimize (weight-0 variations dropped)
    auto live_vars = [&](const XformSpec &x) {
        std::vector<const VarSpec *> out;
        for (const auto &v : x.vars)
            if (std::fabs((R)v.weight) != (R)0.0)
                out.push_back(&v);
        return out;
    };
    std::vector<std::vector<const VarSpec *>> lists;
    std::vector<std::vector<int>> seqs;
    for (int k = 0; k < d.nxf; ++k) {
        lists.push_back(live_vars(spec.xforms[(size_t)d.order[(size_t)k]]));
        std::vector<int> s;
        for (auto *v : lists.back())
            s.push_back(v->type);
        seqs.push_back(s);
    }
    std::vector<int> slots;
    for (const auto &s : seqs)
        slots = scs_merge(slots, s);
    // embed each xform's sequence into the slot sequence (leftmost match)
    const int ns = (int)slots.size();
    std::vector<std::vector<int>> pos(d.nxf, std::vector<int>(ns, -1)); // slot -> var index
    for (int k = 0; k < d.nxf; ++k) {
        size_t j = 0;
        for (int s = 0; s < ns && j < seqs[(size_t)k].size(); ++s)
            if (slots[(size_t)s] == seqs[(size_t)k][j]) {
                pos[(size_t)k][(size_t)s] = (int)j;
                ++j;
            }
        if (j != seqs[(size_t)k].size())
            throw JsonError("internal error: variation slot embedding failed");
    }

    // table layout
    d.lay = xform_layout(dims, r);
    std::vector<R> &T = d.table;
    auto align2 = [](size_t n) { return (n + 1) & ~(size_t)1; };
    T.assign(align2((size_t)d.nxf), (R)0);
    for (int k = 0; k < d.nxf; ++k)
        T[(size_t)k] = d.cw[(size_t)k];
    auto put_xform = [&](const XformSpec &x) {
        size_t base = T.size();
        T.resize(base + (size_t)d.lay.stride, (R)0);
        R *X = T.data() + base;
        for (int i = 0; i < dims * dims; ++i) {
            X[d.lay.preA + i] = (R)x.preA[i];
            X[d.lay.postA + i] = (R)x.postA[i];
        }
        for (int i = 0; i < dims; ++i) {
            X[d.lay.preb + i] = (R)x.preb[i];
            X[d.lay.postb + i] = (R)x.postb[i];
        }
        bool hc = r > 0 && x.has_color;
        X[d.lay.has_color] = hc ? (R)1 : (R)0;
        X[d.lay.has_pre] = x.has_pre ? (R)1 : (R)0;
        X[d.lay.has_post] = x.has_post ? (R)1 : (R)0;
        X[d.lay.speed] = (r > 0 && x.has_speed) ? (R)x.speed : (R)spec.color_speed;
        if (hc)
            for (int i = 0; i < r; ++i)
                X[d.lay.color + i] = (R)x.color[(size_t)i];
        return base;
    };
    d.xf0 = (int)T.size();
    for (int k = 0; k < d.nxf; ++k)
        put_xform(spec.xforms[(size_t)d.order[(size_t)k]]);
    d.fin0 = (int)T.size();
    if (spec.has_final)
        put_xform(spec.final_xf);
    // slot entries: [weight, params..., axis_x, axis_y] per xform
    auto entry_size = [&](int type) {
        return 1 + var_nparams(type, dims) + ((dims == 3 && type >= kNumNdVariations) ? 2 : 0);
    };
    for (int s = 0; s < ns; ++s) {
        int type = slots[(size_t)s];
        int es = entry_size(type);
        d.slot_type.push_back(type);
        d.slot_off.push_back((int)T.size());
        d.slot_stride.push_back(es);
        size_t base = T.size();
        T.resize(base + (size_t)(es * d.nxf), (R)0);
        for (int k = 0; k < d.nxf; ++k) {
            int j = pos[(size_t)k][(size_t)s];
            if (j < 0)
                continue;
            const VarSpec &v = *lists[(size_t)k][(size_t)j];
            R *E = T.data() + base + (size_t)(k * es);
            E[0] = (R)v.weight;
            R tmp[24] = {};
            var_setup<R>(v, dims, tmp);
            int np = var_nparams(type, dims);
            int pofs = 1;
            if (dims == 3 && type >= kNumNdVariations) {
                E[1] = (R)v.axis_x;
                E[2] = (R)v.axis_y;
                pofs = 3;
            }
            for (int i = 0; i < np; ++i)
                E[pofs + i] = tmp[i];
        }
    }
    if (spec.has_final) {
        for (const VarSpec *v : live_vars(spec.final_xf)) {
            int es = entry_size(v->type);
            d.fin_type.push_back(v->type);
            d.fin_off.push_back((int)T.size());
            size_t base = T.size();
            T.resize(base + (size_t)es, (R)0);
            R *E = T.data() + base;
            E[0] = (R)v->weight;
            R tmp[24] = {};
            var_setup<R>(*v, dims, tmp);
            int np = var_nparams(v->type, dims);
            int pofs = 1;
            if (dims == 3 && v->type >= kNumNdVariations) {
                E[1] = (R)v->axis_x;
                E[2] = (R)v->axis_y;
                pofs = 3;
            }
            for (int i = 0; i < np; ++i)
                E[pofs + i] = tmp[i];
        }
    }
    // per-xform variation lists (see DeviceFlame::bt0)
    {
        auto code_of = [](int type) {
            switch (type) {
            case 0: case 1: case 2: return type; // linear, sinusoidal, spherical
            case 20: case 21: case 22: return type - 17; // swirl, horseshoe, polar
            default: return -1;
            }
        };
        bool basic = dims == 2;
        int nvmax = 0;
        for (int k = 0; k < d.nxf && basic; ++k) {
            nvmax = std::max(nvmax, (int)lists[(size_t)k].size());
            for (const VarSpec *v : lists[(size_t)k])
                basic = basic && code_of(v->type) >= 0;
        }
        if (basic) {
            d.bt_nvmax = nvmax;
            d.bt_stride = 2 + 2 * nvmax;
            d.bt0 = (int)T.size();
            T.resize(T.size() + (size_t)(d.bt_stride * d.nxf), (R)0);
            for (int k = 0; k < d.nxf; ++k) {
                R *B = T.data() + d.bt0 + (size_t)(k * d.bt_stride);
                int mask = 0;
                const auto &L = lists[(size_t)k];
                B[0] = (R)(int)L.size();
                for (size_t j = 0; j < L.size(); ++j) {
                    const int c = code_of(L[j]->type);
                    mask |= 1 << c;
                    B[2 + 2 * j] = (R)c;
                    B[3 + 2 * j] = (R)L[j]->weight;
                }
                B[1] = (R)mask;
            }
        }
    }
    T.resize(align2(T.size()), (R)0);
    if (ns > 64)
        throw JsonError("flame has " + std::to_string(ns) +
                        " variation slots after merging the xforms' variation orders "
                        "(at most 64 are supported)");
    d.masks.assign((size_t)d.nxf, 0ull);
    for (int k = 0; k < d.nxf; ++k)
        for (int s = 0; s < ns; ++s)
            if (pos[(size_t)k][(size_t)s] >= 0)
                d.masks[(size_t)k] |= 1ull << s;
    for (int t : slots)
        d.type_set[t / 32] |= 1u << (t % 32);
    for (int t : d.fin_type)
        d.type_set[t / 32] |= 1u << (t % 32);
    for (int k = 0; k < d.nxf; ++k) {
        int dr = 0;
        for (auto *v : lists[(size_t)k])
            dr += variation_draws(v->type, dims);
        d.max_draws = std::max(d.max_draws, dr);
    }
    return d;
}

template DeviceFlame<double> compile_flame<double>(const FlameSpec &);
template DeviceFlame<float> compile_flame<float>(const FlameSpec &);

} // namespace ffr

// Flame description read from JSON and its compiled device form.
//
// FlameSpec mirrors what the reference's Flame<dims>(Json) accepts and
// rejects (src/types/flame.hpp:91-210, src/types/xform.hpp:71-172,
// src/types/affine.hpp:45-92, src/types/point.hpp:29-47, variation
// constructors in src/variations/variations.hpp), with its error texts.
//
// DeviceFlame<R> is the flame after the reference's optimisation steps
// (drop weight-0 xforms and variations, std::sort by decreasing weight,
// normalised cumulative weights; flame.hpp:44-83, xform.hpp:47-58) laid out
// as one flat table of R values that the kernels copy into LDS.
#pragma once

#include "json.hpp"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace ffr {

constexpr int kNumVariations = 98;
constexpr int kNumNdVariations = 20; // ids 0-19 are N-D, 20-97 are 2-D
constexpr int kMaxColorDims = 127;
constexpr uint64_t kMaxDim = 65535;

struct VariationInfo {
    const char *name;
    std::vector<std::pair<const char *, char>> keys; // 's' scalar, 'p' point, '2' 2-D point
};

const VariationInfo &variation_info(int type);
int variation_id(const std::string &name); // -1 when unknown
// words drawn from the walker stream by one evaluation (fixed per type)
int variation_draws(int type, int dims);

struct VarSpec {
    int type = 0;
    double weight = 0;
    std::vector<double> params; // raw values, catalogue key order
    int axis_x = -1, axis_y = -1;
};

struct XformSpec {
    int id = -1;
    double weight = 1.0;
    bool has_pre = false, has_post = false;
    double preA[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double preb[3] = {0, 0, 0};
    double postA[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double postb[3] = {0, 0, 0};
    bool has_color = false;
    std::vector<double> color;
    bool has_speed = false;
    double speed = 0.5;
    std::vector<VarSpec> vars;
};

struct FlameSpec {
    int dims = 2;
    uint64_t size[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    int color_dims = 0;
    double color_speed = 0.5;
    std::vector<XformSpec> xforms; // file order
    bool has_final = false;
    XformSpec final_xf;
    uint64_t cells() const;
};

// dims as the CLI reads it (ffr_buf.cpp:131); throws JsonError
int flame_dims(const Json &doc);
// full validation for dims in {1,2,3}; fp32 = compare in float like the
// reference's float build
FlameSpec parse_flame(const Json &doc, bool fp32);

// ---------------------------------------------------------------------
// compiled form

// offsets inside one xform record of the table (in R units)
struct XformLayout {
    int preA, preb, postA, postb, speed, color, has_color, has_pre, has_post;
    int stride;
};
XformLayout xform_layout(int dims, int color_dims);

template <typename R> struct DeviceFlame {
    int dims = 2, r = 0;
    int nxf = 0, n_ids = 0;
    int has_final = 0;
    std::vector<int> order;    // sorted position -> file index (xform id)
    std::vector<R> cw;         // cumulative weights
    R lo[3]{}, hi[3]{}, mult_d[3]{};
    uint64_t mult_i[3]{};
    uint64_t cells = 1;
    XformLayout lay{};
    // flat table (copied to LDS): [cw | xform records | final record |
    // slot entries | final variation entries]
    std::vector<R> table;
    int xf0 = 0, fin0 = 0;
    // variation slots: a common supersequence of every xform's variation
    // type sequence, so summing a lane's used slots in slot order is its
    // xform's JSON order
    std::vector<int> slot_type, slot_off, slot_stride;
    std::vector<unsigned long long> masks; // [nxf] bit s set when the xform uses slot s
    std::vector<int> fin_type, fin_off;
    // per-xform variation lists for 2-D flames whose variations are all of
    // the basic six (linear, sinusoidal, spherical, swirl, horseshoe,
    // polar; codes 0..5): [count, code mask, (code, weight) x bt_nvmax]
    // per xform in JSON order, at table offset bt0 (-1: not built)
    int bt0 = -1, bt_stride = 0, bt_nvmax = 0;
    uint32_t type_set[4] = {0, 0, 0, 0}; // bitset of variation types present
    int max_draws = 0;
};

template <typename R>
DeviceFlame<R> compile_flame(const FlameSpec &spec); // throws JsonError

// the reference's BufferRenderer limits
constexpr uint64_t kMaxCells = 1ull << 48;

} // namespace ffr

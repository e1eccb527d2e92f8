// Per-flame compiled walker kernels (rtc.cpp).
#pragma once

#include "walk.hpp"

#include <string>
#include <vector>

namespace ffr {

// the instantiation of the walker kernel a flame's own kernel compiles:
// precision mode, dimensions, colour registers (0, 4, 128), variation set,
// LDS cell cache, and the most lanes per workgroup it is built for
struct RtcKernel {
    int mode = 0, dims = 2, nc = 0, vs = 0;
    bool cache = false;
    int block_max = 256;
    bool bin = false; // a pass of binned accumulation (walk_bin_tasks)
};

// the flame's kernel (a hipFunction_t) for this flame (its arguments and
// cumulative weights) and instantiation,
// compiled on first use and kept for the process; compile_s receives the
// time the compile took (0 after the first call)
template <typename R>
int rtc_flame_kernel(const WalkArgs<R> &a, const std::vector<R> &cw, const RtcKernel &k,
                     int device, void **fn, double *compile_s, std::string &err);

// the generated source, and hiprtc alone (no device needed): the code
// object for `arch`, or the compiler's log in err
std::string rtc_flame_source_f64(const WalkArgs<double> &a, const std::vector<double> &cw,
                                 const RtcKernel &k);
std::string rtc_flame_source_f32(const WalkArgs<float> &a, const std::vector<float> &cw,
                                 const RtcKernel &k);
int rtc_compile(const std::string &src, const std::string &arch, std::vector<char> &code,
                std::string &err);

} // namespace ffr

// Per-flame compiled walker kernels.
//
// The reference evaluates a flame through objects: Flame::getRandomXForm
// scans the cumulative weights, XForm::applyIteration loops over shared
// pointers to Variation objects and calls a virtual calc for each
// (src/types/flame.hpp:212-219, src/types/xform.hpp:211-227,
// src/variations/variations.hpp:38-59).  The generic walker kernel does the
// same work from a table: the xform count, the table layout, the variation
// slots and their types, the bounds and cell geometry are kernel arguments,
// read at run time.  Here a flame gets its own kernel instead: the same
// walker code (walk_kernel.inc) compiled with hiprtc for a type whose flame
// fields are compile-time constants, so loops over xforms and slots have
// constant trip counts, the switch over variation types folds to the
// flame's own variations and the bounds are literals.  Per-xform
// coefficients stay in the LDS table, indexed by each lane's xform.
//
// Code objects are kept per process, keyed by the generated source (which
// holds every constant) and the device's architecture.
#include "rtc.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

// the walker sources, embedded in the library when it is built
asm(".pushsection .rodata\n"
    ".global ffr_rtc_src_walk_kernel\n"
    "ffr_rtc_src_walk_kernel:\n"
    ".incbin \"csrc/walk_kernel.inc\"\n"
    ".byte 0\n"
    ".global ffr_rtc_src_variations\n"
    "ffr_rtc_src_variations:\n"
    ".incbin \"csrc/variations_dev.hpp\"\n"
    ".byte 0\n"
    ".global ffr_rtc_src_device_common\n"
    "ffr_rtc_src_device_common:\n"
    ".incbin \"csrc/device_common.hpp\"\n"
    ".byte 0\n"
    ".global ffr_rtc_src_walk\n"
    "ffr_rtc_src_walk:\n"
    ".incbin \"csrc/walk.hpp\"\n"
    ".byte 0\n"
    ".popsection\n");
extern "C" const char ffr_rtc_src_walk_kernel[];
extern "C" const char ffr_rtc_src_variations[];
extern "C" const char ffr_rtc_src_device_common[];
extern "C" const char ffr_rtc_src_walk[];

namespace ffr {

namespace {

struct Compiled {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    double compile_s = 0;
};

std::mutex g_mu;
std::unordered_map<std::string, Compiled> g_cache;

// exact literals: hexadecimal floating point
std::string lit(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%a", v);
    return b;
}
std::string lit(float v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%af", (double)v);
    return b;
}

template <typename T> std::string array_lit(const T *v, int n)
{
    std::ostringstream o;
    o << "{";
    for (int i = 0; i < n; ++i)
        o << (i ? ", " : "") << v[i];
    o << "}";
    return o.str();
}

template <typename R> std::string real_array(const R *v, int n)
{
    std::ostringstream o;
    o << "{";
    if (n == 0)
        o << "0";
    for (int i = 0; i < n; ++i)
        o << (i ? ", " : "") << lit(v[i]);
    o << "}";
    return o.str();
}

const char *mode_types(int mode)
{
    switch (mode) {
    case MODE_F64_U64: return "double, unsigned long long, double";
    case MODE_F32_U32: return "float, unsigned int, float";
    default: return "float, unsigned long long, double";
    }
}

// trajectory and count types of the binned walker (count-only buffers)
const char *bin_types(int mode)
{
    switch (mode) {
    case MODE_F64_U64: return "double, unsigned long long";
    case MODE_F32_U32: return "float, unsigned int";
    default: return "float, unsigned long long";
    }
}

template <typename R>
std::string flame_source(const WalkArgs<R> &a, const std::vector<R> &cw, const RtcKernel &k)
{
    const char *rt = sizeof(R) == 8 ? "double" : "float";
    std::ostringstream s;
    s << "#include \"walk_kernel.inc\"\n"
         "namespace ffr {\n"
         "struct FlameArgs : WalkDyn<"
      << rt << "> {\n";
    s << "    static constexpr bool kCompiled = true;\n";
    s << "    static constexpr " << rt << " lo[3] = " << real_array(a.lo, 3) << ";\n";
    s << "    static constexpr " << rt << " hi[3] = " << real_array(a.hi, 3) << ";\n";
    s << "    static constexpr " << rt << " mult_d[3] = " << real_array(a.mult_d, 3) << ";\n";
    s << "    static constexpr uint64_t mult_i[3] = {" << a.mult_i[0] << "ull, " << a.mult_i[1]
      << "ull, " << a.mult_i[2] << "ull};\n";
    s << "    static constexpr uint64_t cells = " << a.cells << "ull;\n";
    s << "    static constexpr int nxf = " << a.nxf << ", n_slots = " << a.n_slots
      << ", has_final = " << a.has_final << ", fin_nv = " << a.fin_nv << ", r = " << a.r
      << ";\n";
    s << "    static constexpr int tab_reals = " << a.tab_reals << ", xf0 = " << a.xf0
      << ", fin0 = " << a.fin0 << ";\n";
    const int ns = a.n_slots > 0 ? a.n_slots : 1;
    const int nf = a.fin_nv > 0 ? a.fin_nv : 1;
    s << "    static constexpr int slot_type[" << ns << "] = " << array_lit(a.slot_type, ns)
      << ";\n";
    s << "    static constexpr int slot_off[" << ns << "] = " << array_lit(a.slot_off, ns)
      << ";\n";
    s << "    static constexpr int slot_stride[" << ns << "] = " << array_lit(a.slot_stride, ns)
      << ";\n";
    s << "    static constexpr int fin_type[" << nf << "] = " << array_lit(a.fin_type, nf)
      << ";\n";
    s << "    static constexpr int fin_off[" << nf << "] = " << array_lit(a.fin_off, nf) << ";\n";
    s << "    static constexpr int bt0 = " << a.bt0 << ", bt_stride = " << a.bt_stride
      << ", bt_nvmax = " << a.bt_nvmax << ", bt_union = " << a.bt_union << ";\n";
    s << "    static constexpr " << rt << " cw[" << (cw.empty() ? 1 : cw.size())
      << "] = " << real_array(cw.data(), (int)cw.size()) << ";\n";
    s << "    static constexpr int l_preA = " << a.l_preA << ", l_preb = " << a.l_preb
      << ", l_postA = " << a.l_postA << ", l_postb = " << a.l_postb
      << ", l_speed = " << a.l_speed << ", l_color = " << a.l_color
      << ", l_has_color = " << a.l_has_color << ", l_has_pre = " << a.l_has_pre
      << ", l_has_post = " << a.l_has_post << ", l_stride = " << a.l_stride << ";\n";
    s << "};\n} // namespace ffr\n";
    s << "extern \"C\" __global__ void __launch_bounds__(" << k.block_max
      << ") ffr_flame_walk(const ffr::FlameArgs *__restrict__ args)\n{\n"
         "    extern __shared__ __align__(16) unsigned char smem[];\n";
    if (k.bin)
        s << "    ffr::walk_bin_tasks<" << bin_types(k.mode) << ", " << k.dims << ", "
          << (k.nc > 0 ? 4 : 0) << ", " << k.vs << ">(*args, smem);\n}\n";
    else
        s << "    ffr::walk_tasks<" << mode_types(k.mode) << ", " << k.dims << ", " << k.nc << ", "
          << k.vs << ", " << (k.cache ? "true" : "false") << ">(*args, smem);\n}\n";
    return s.str();
}

int load(const std::vector<char> &code, Compiled &out, std::string &log)
{
    hipError_t e = hipModuleLoadData(&out.module, code.data());
    if (e == hipSuccess)
        e = hipModuleGetFunction(&out.fn, out.module, "ffr_flame_walk");
    if (e != hipSuccess) {
        log = std::string("loading the compiled flame kernel: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

} // namespace

int rtc_compile(const std::string &src, const std::string &arch, std::vector<char> &code,
                std::string &log)
{
    const char *headers[] = {ffr_rtc_src_walk_kernel, ffr_rtc_src_variations,
                             ffr_rtc_src_device_common, ffr_rtc_src_walk};
    const char *names[] = {"walk_kernel.inc", "variations_dev.hpp", "device_common.hpp",
                           "walk.hpp"};
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "flame_walk.hip", 4, headers, names) !=
        HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return -1;
    }
    // the library's own build flags (Makefile HIPFLAGS): no FMA
    // contraction, correctly rounded float division and square root,
    // denormals kept, no packed FP32 instructions
    const std::string off = "--offload-arch=" + arch;
    const char *opts[] = {off.c_str(),
                          "-O3",
                          "-std=c++17",
                          "-ffp-contract=off",
                          "-fno-gpu-flush-denormals-to-zero",
                          "-fhip-fp32-correctly-rounded-divide-sqrt",
                          "-Xclang",
                          "-target-feature",
                          "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    log.assign(n, '\0');
    if (n)
        hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return -1;
    }
    size_t bytes = 0;
    hiprtcGetCodeSize(prog, &bytes);
    code.resize(bytes);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return 0;
}

template <typename R>
int rtc_flame_kernel(const WalkArgs<R> &a, const std::vector<R> &cw, const RtcKernel &k,
                     int device, void **fn, double *compile_s, std::string &err)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        return -1;
    }
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos)
        arch.resize(colon);
    const std::string src = flame_source(a, cw, k);
    const std::string key = arch + "|" + std::to_string(device) + "|" + src;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) {
        // FFR_RTC_DUMP: a directory that receives each generated source
        // (to read or rebuild a flame's kernel offline)
        if (const char *dir = std::getenv("FFR_RTC_DUMP")) {
            const std::string path = std::string(dir) + "/ffr_flame_" +
                                     std::to_string(std::hash<std::string>{}(src)) + ".hip";
            if (FILE *f = std::fopen(path.c_str(), "w")) {
                std::fputs(src.c_str(), f);
                std::fclose(f);
            }
        }
        Compiled c;
        std::string log;
        std::vector<char> code;
        auto t0 = std::chrono::steady_clock::now();
        if (rtc_compile(src, arch, code, log) != 0) {
            err = "compiling the flame's kernel: " + log;
            return -1;
        }
        if (load(code, c, log) != 0) {
            err = log;
            return -1;
        }
        c.compile_s =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        it = g_cache.emplace(key, c).first;
    }
    *fn = (void *)it->second.fn;
    if (compile_s)
        *compile_s = it->second.compile_s;
    return 0;
}

template int rtc_flame_kernel<double>(const WalkArgs<double> &, const std::vector<double> &,
                                      const RtcKernel &, int, void **, double *, std::string &);
template int rtc_flame_kernel<float>(const WalkArgs<float> &, const std::vector<float> &,
                                     const RtcKernel &, int, void **, double *, std::string &);

std::string rtc_flame_source_f64(const WalkArgs<double> &a, const std::vector<double> &cw,
                                 const RtcKernel &k)
{
    return flame_source(a, cw, k);
}
std::string rtc_flame_source_f32(const WalkArgs<float> &a, const std::vector<float> &cw,
                                 const RtcKernel &k)
{
    return flame_source(a, cw, k);
}

} // namespace ffr

// ffr-buf: drop-in for the reference's buffer renderer program
// (src/ffr_buf.cpp:61-274).  Same options (-h -f -i -o -s -t -b -z), the
// same log lines on stderr, the same raw buffer on the output; the render
// runs on the GPU through libffr_hip.
//
// Additions (all optional): --seed (walker stream key), --precision
// (fp64 | fp32), --word (64 | 32), --settle (settle iterations),
// --device (HIP device).  -t is accepted and reported; the GPU ignores it.
#include "../../include/ffr.h"

#include "json.hpp"

#include <unistd.h>

#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

const char *kVersion = "unspecified";

struct Args {
    std::string flame, output;
    std::vector<std::string> inputs;
    uint64_t samples = 0;
    uint64_t threads = 0;
    uint64_t batch = 0;
    uint64_t bad = 256;
    uint64_t seed = 0;
    std::string precision = "fp64";
    int word = 0; // 0: default for the precision
    int32_t settle = -1;
    int device = 0;
    bool help = false;
    bool have_flame = false, have_output = false;
};

void usage(std::ostream &os)
{
    os << "ffr-buf usage:\n"
          "  -h [ --help ]                 show this help message\n"
          "  -f [ --flame ] arg            flame parameters JSON file\n"
          "  -i [ --input ] arg            buffers to add initially\n"
          "  -o [ --output ] arg           output file\n"
          "  -s [ --samples ] arg (=0)     samples to render\n"
          "  -t [ --threads ] arg          threads to use (accepted; the GPU ignores it)\n"
          "  -b [ --batch-size ] arg (=0)  batch size (samples per walker)\n"
          "  -z [ --bad-values ] arg (=256) bad values limit\n"
          "  --seed arg (=0)               key of the walkers' random streams\n"
          "  --precision arg (=fp64)       trajectory precision: fp64 or fp32\n"
          "  --word arg                    buffer word bits: 64 (default) or 32\n"
          "  --settle arg                  settle iterations (default 53 for fp64, 24 for fp32)\n"
          "  --device arg (=0)             HIP device\n";
}

bool parse_u64(const std::string &s, uint64_t &v)
{
    if (s.empty() || s[0] == '-')
        return false;
    errno = 0;
    char *end = nullptr;
    unsigned long long x = std::strtoull(s.c_str(), &end, 0);
    if (errno || *end)
        return false;
    v = x;
    return true;
}

// boost::program_options style: -f x, -fx, --flame x, --flame=x
int parse_args(int argc, char **argv, Args &a)
{
    struct Opt {
        const char *lng;
        char shrt;
        bool takes;
    };
    static const Opt opts[] = {
        {"help", 'h', false},      {"flame", 'f', true},     {"input", 'i', true},
        {"output", 'o', true},     {"samples", 's', true},   {"threads", 't', true},
        {"batch-size", 'b', true}, {"bad-values", 'z', true}, {"seed", 0, true},
        {"precision", 0, true},    {"word", 0, true},        {"settle", 0, true},
        {"device", 0, true},
    };
    for (int i = 1; i < argc; ++i) {
        std::string tok = argv[i];
        const Opt *o = nullptr;
        std::string val;
        bool have_val = false;
        if (tok.size() > 2 && tok[0] == '-' && tok[1] == '-') {
            std::string name = tok.substr(2);
            size_t eq = name.find('=');
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
                have_val = true;
            }
            for (const auto &c : opts)
                if (name == c.lng)
                    o = &c;
        } else if (tok.size() >= 2 && tok[0] == '-' && tok[1] != '-') {
            for (const auto &c : opts)
                if (c.shrt && tok[1] == c.shrt)
                    o = &c;
            if (o && tok.size() > 2) {
                val = tok.substr(2);
                have_val = true;
            }
        }
        if (!o) {
            std::cerr << "ERROR: unrecognised option '" << tok << "'" << std::endl;
            return 1;
        }
        if (o->takes && !have_val) {
            if (i + 1 >= argc) {
                std::cerr << "ERROR: the required argument for option '--" << o->lng
                          << "' is missing" << std::endl;
                return 1;
            }
            val = argv[++i];
        }
        std::string n = o->lng;
        uint64_t u = 0;
        auto need_u64 = [&](uint64_t &dst) {
            if (!parse_u64(val, u)) {
                std::cerr << "ERROR: the argument ('" << val << "') for option '--" << n
                          << "' is invalid" << std::endl;
                return false;
            }
            dst = u;
            return true;
        };
        if (n == "help") {
            a.help = true;
        } else if (n == "flame") {
            a.flame = val;
            a.have_flame = true;
        } else if (n == "input") {
            a.inputs.push_back(val);
        } else if (n == "output") {
            a.output = val;
            a.have_output = true;
        } else if (n == "samples") {
            if (!need_u64(a.samples))
                return 1;
        } else if (n == "threads") {
            if (!need_u64(a.threads))
                return 1;
        } else if (n == "batch-size") {
            if (!need_u64(a.batch))
                return 1;
        } else if (n == "bad-values") {
            if (!need_u64(a.bad))
                return 1;
        } else if (n == "seed") {
            if (!need_u64(a.seed))
                return 1;
        } else if (n == "precision") {
            if (val != "fp64" && val != "fp32") {
                std::cerr << "ERROR: --precision must be fp64 or fp32" << std::endl;
                return 1;
            }
            a.precision = val;
        } else if (n == "word") {
            if (val != "64" && val != "32") {
                std::cerr << "ERROR: --word must be 64 or 32" << std::endl;
                return 1;
            }
            a.word = std::atoi(val.c_str());
        } else if (n == "settle") {
            uint64_t s = 0;
            if (!need_u64(s))
                return 1;
            a.settle = (int32_t)s;
        } else if (n == "device") {
            uint64_t d = 0;
            if (!need_u64(d))
                return 1;
            a.device = (int)d;
        }
    }
    return 0;
}

bool read_all(std::istream &is, std::string &out)
{
    out.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
    return !is.bad();
}

// stream callbacks of ffr_add_buffer_stream / ffr_read_buffer_stream
size_t istream_read(void *ctx, void *dst, size_t n)
{
    auto *is = static_cast<std::istream *>(ctx);
    is->read(static_cast<char *>(dst), (std::streamsize)n);
    return (size_t)is->gcount();
}

size_t ostream_write(void *ctx, const void *src, size_t n)
{
    auto *os = static_cast<std::ostream *>(ctx);
    os->write(static_cast<const char *>(src), (std::streamsize)n);
    return os->good() ? n : 0;
}

double now_sec()
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

} // namespace

int main(int argc, char **argv)
{
    Args a;
    a.threads = std::thread::hardware_concurrency();
    if (a.threads == 0)
        a.threads = 1;
    if (parse_args(argc, argv, a))
        return 1;
    if (a.help || argc < 2) {
        usage(std::cerr);
        return 1;
    }
    if (!a.have_flame || !a.have_output) {
        std::cerr << "ERROR: the option '--" << (a.have_flame ? "output" : "flame")
                  << "' is required but missing" << std::endl;
        return 1;
    }

    ffr_options opt;
    ffr_default_options(&opt);
    opt.precision = a.precision == "fp32" ? FFR_FP32 : FFR_FP64;
    // the default build's layout unless --word 32 asks for the float build's
    opt.word = a.word == 32 ? FFR_WORD32 : FFR_WORD64;
    opt.settle_iters = a.settle;
    opt.device = a.device;
    opt.seed = a.seed;
    opt.batch_size = a.batch;
    opt.bad_value_limit = a.bad;

    // echo the options (ffr_buf.cpp:104-119)
    std::cerr << "ffr-buf version " << kVersion << std::endl;
    std::cerr << "--flame " << a.flame << std::endl;
    for (auto &s : a.inputs)
        std::cerr << "--input " << s << std::endl;
    std::cerr << "--output " << a.output << std::endl;
    std::cerr << "--samples " << a.samples;
    if (a.samples == 0)
        std::cerr << " (not rendering)";
    std::cerr << std::endl;
    std::cerr << "--threads " << a.threads << std::endl;

    // the flame (ffr_buf.cpp:121-143)
    std::string text;
    if (a.flame == "-") {
        read_all(std::cin, text);
    } else {
        std::ifstream f(a.flame, std::ios::in | std::ios::binary);
        if (!f) {
            std::cerr << "ERROR: cannot open flame file " << a.flame << std::endl;
            return 1;
        }
        read_all(f, text);
    }
    ffr::Json doc;
    int64_t dims = 0;
    try {
        doc = ffr::Json::parse(text);
        dims = doc["dimensions"].intValue();
    } catch (const std::exception &e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 1;
    }

    ffr_renderer *r = nullptr;
    const bool dims_ok = dims >= 1 && dims <= 3;
    int rc = FFR_OK;
    if (dims_ok) {
        rc = ffr_create(text.data(), text.size(), &opt, &r);
    }
    uint64_t batch = a.batch;
    bool calculated = false;
    if (batch == 0) {
        batch = r ? ffr_auto_batch(r, a.samples) : 0;
        calculated = true;
    }
    if (calculated)
        std::cerr << "--batch_size 0 (" << batch << ")" << std::endl;
    else
        std::cerr << "--batch_size " << batch << std::endl;
    std::cerr << "--bad_values " << a.bad << std::endl;
    std::cerr << "--" << std::endl;
    std::cerr << "flame: " << doc.dump() << std::endl;
    if (!dims_ok) {
        std::cerr << "ERROR: " << dims << "D not supported" << std::endl;
        return 1;
    }
    if (rc != FFR_OK) {
        std::cerr << "ERROR: " << ffr_last_error() << std::endl;
        return 1;
    }
    if (a.samples > 0 && a.batch != 0 && a.batch < 256) {
        // BufferRenderer::render (buffer_renderer.hpp:283-285)
        std::cerr << "ERROR: BufferRenderer::render(): batch size too small" << std::endl;
        ffr_destroy(r);
        return 1;
    }

    const size_t bytes = ffr_buffer_bytes(r);
    std::cerr << "buffer: " << bytes << " bytes, little endian, " << ffr_word_bytes(r)
              << " byte numbers" << std::endl;
    std::cerr << "size:";
    {
        const auto &sz = doc["size"].arrayValue();
        for (const auto &s : sz)
            std::cerr << " " << (uint64_t)s.floatValue();
    }
    std::cerr << std::endl;
    std::cerr << "color: " << ffr_color_dims(r) << " dimensions" << std::endl;

    // add the input buffers (ffr_buf.cpp:167-183), streamed to the device
    for (auto &s : a.inputs) {
        std::cerr << "adding input file " << s << std::endl;
        std::ifstream f;
        std::istream *is = &std::cin;
        if (s != "-") {
            f.open(s, std::ios::in | std::ios::binary);
            is = &f;
        }
        if (ffr_add_buffer_stream(r, istream_read, is) != FFR_OK) {
            std::cerr << "ERROR: " << ffr_last_error() << std::endl;
            ffr_destroy(r);
            return 1;
        }
    }

    // render (ffr_buf.cpp:185-257)
    if (a.samples > 0) {
        std::cerr << "render start" << std::endl;
        double t1 = now_sec();
        ffr_stats st;
        rc = ffr_render(r, a.samples, &st);
        double t2 = now_sec();
        if (rc != FFR_OK && rc != FFR_E_BADVALUE_LIMIT) {
            std::cerr << "ERROR: " << ffr_last_error() << std::endl;
            ffr_destroy(r);
            return 1;
        }
        std::cerr << "progress: " << st.walkers << "/" << st.walkers << " batches, "
                  << a.samples << "/" << a.samples << " samples (100%) " << (t2 - t1)
                  << "sec elapsed, (~0sec remaining)";
        std::cerr << std::endl;
        double secs = t2 - t1;
        if (secs < 1e-9)
            secs = 1e-9;
        std::cerr << "render done: " << secs << " sec ("
                  << (uint64_t)((double)a.samples / secs) << " samples/sec)" << std::endl;
        if (rc == FFR_E_BADVALUE_LIMIT)
            std::cerr << "render failure "
                      << " (flame may not be sufficiently contractive)" << std::endl;
        double iter_part = (double)st.s_iter / (double)a.samples;
        std::cerr << "samples iterated: " << st.s_iter << " (" << (100 * iter_part) << "%)"
                  << std::endl;
        double plot_part = (double)st.s_plot / (double)a.samples;
        std::cerr << "samples plotted: " << st.s_plot << " (" << (100 * plot_part) << "%)"
                  << std::endl;
        std::vector<uint64_t> dist(ffr_xform_id_count(r));
        ffr_xform_distribution(r, dist.data(), dist.size());
        std::cerr << "xform selection:";
        for (auto n : dist)
            std::cerr << " " << n;
        std::cerr << std::endl;
        size_t nbad = 0;
        ffr_bad_values(r, nullptr, nullptr, 0, &nbad);
        std::vector<uint64_t> ids(nbad);
        std::vector<double> pts(nbad * (size_t)dims);
        ffr_bad_values(r, ids.data(), pts.data(), nbad, &nbad);
        std::cerr << "bad value xforms:";
        for (auto n : ids)
            std::cerr << " " << n;
        std::cerr << std::endl;
        std::cerr << "bad value points:";
        for (size_t i = 0; i < ids.size(); ++i) {
            std::cerr << " (" << pts[i * (size_t)dims];
            for (int k = 1; k < dims; ++k)
                std::cerr << "," << pts[i * (size_t)dims + (size_t)k];
            std::cerr << ")";
        }
        std::cerr << std::endl;
        std::cerr << "extreme coordinates:";
        for (int k = 0; k < dims; ++k)
            std::cerr << " (" << st.pt_min[k] << "," << st.pt_max[k] << ")";
        std::cerr << std::endl;
    }

    // write the buffer (ffr_buf.cpp:259-272), streamed from the device
    std::cerr << "writing output" << std::endl;
    std::ofstream f;
    std::ostream *os = &std::cout;
    if (a.output != "-") {
        f.open(a.output, std::ios::out | std::ios::binary);
        os = &f;
    }
    int wrc = os->good() ? ffr_read_buffer_stream(r, ostream_write, os) : FFR_E_IO;
    ffr_destroy(r);
    if (wrc == FFR_OK) {
        os->flush();
        if (!os->good())
            wrc = FFR_E_IO;
    }
    if (wrc != FFR_OK) {
        std::cerr << "ERROR: error writing output" << std::endl;
        return 1;
    }
    return 0;
}

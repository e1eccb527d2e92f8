// ffr-img: drop-in for the reference's image program (src/ffr_img.cpp:60-309).
// Same options (-h -f -i -o -y -m -g -c -b), the same log lines on stderr,
// the same pixels; the input buffers are summed and tonemapped on the GPU
// through libffr_hip, and the PNG is written here with zlib.
//
// Additions (all optional): --word (64 | 32: the buffer layout of the
// reference's default or float build), --device (HIP device).
#include "../../include/ffr.h"

#include "json.hpp"

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

namespace {

const char *kVersion = "unspecified";

struct Args {
    std::string flame, output;
    std::vector<std::string> inputs;
    double gamma = 1.0;
    std::string gamma_text; // parsed again as a float for --word 32
    uint64_t bits = 8;
    bool mono = false, gray = false, color = false;
    int word = 64;
    int device = 0;
    bool help = false;
    bool have_flame = false, have_output = false;
};

void usage(std::ostream &os)
{
    os << "ffr-img usage:\n"
          "  -h [ --help ]            show this help message\n"
          "  -f [ --flame ] arg       flame parameters JSON file\n"
          "  -i [ --input ] arg       buffers to add\n"
          "  -o [ --output ] arg      output file\n"
          "  -y [ --gamma ] arg (=1)  gamma value\n"
          "  -m [ --monochrome ]      binary image from histogram\n"
          "  -g [ --grayscale ]       grayscale image from histogram\n"
          "  -c [ --color ]           color image from 3d color only\n"
          "  -b [ --bits ] arg (=8)   bits per color channel\n"
          "  --word arg (=64)         buffer word bits: 64 or 32 (float build layout)\n"
          "  --device arg (=0)        HIP device\n";
}

bool parse_u64(const std::string &s, uint64_t &v)
{
    if (s.empty() || s[0] == '-')
        return false;
    errno = 0;
    char *end = nullptr;
    unsigned long long x = std::strtoull(s.c_str(), &end, 10);
    if (errno || *end)
        return false;
    v = x;
    return true;
}

bool parse_double(const std::string &s, double &v)
{
    if (s.empty())
        return false;
    errno = 0;
    char *end = nullptr;
    double x = std::strtod(s.c_str(), &end);
    if (errno || *end)
        return false;
    v = x;
    return true;
}

// boost::program_options style: -f x, -fx, --flame x, --flame=x; the three
// colouring switches take no value
int parse_args(int argc, char **argv, Args &a)
{
    struct Opt {
        const char *lng;
        char shrt;
        bool takes;
    };
    static const Opt opts[] = {
        {"help", 'h', false},      {"flame", 'f', true},  {"input", 'i', true},
        {"output", 'o', true},     {"gamma", 'y', true},  {"monochrome", 'm', false},
        {"grayscale", 'g', false}, {"color", 'c', false}, {"bits", 'b', true},
        {"word", 0, true},         {"device", 0, true},
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
        if (!o || (!o->takes && have_val)) {
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
        const std::string n = o->lng;
        auto invalid = [&]() {
            std::cerr << "ERROR: the argument ('" << val << "') for option '--" << n
                      << "' is invalid" << std::endl;
            return 1;
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
        } else if (n == "gamma") {
            if (!parse_double(val, a.gamma))
                return invalid();
            a.gamma_text = val;
        } else if (n == "monochrome") {
            a.mono = true;
        } else if (n == "grayscale") {
            a.gray = true;
        } else if (n == "color") {
            a.color = true;
        } else if (n == "bits") {
            if (!parse_u64(val, a.bits))
                return invalid();
        } else if (n == "word") {
            if (val != "64" && val != "32") {
                std::cerr << "ERROR: --word must be 64 or 32" << std::endl;
                return 1;
            }
            a.word = std::atoi(val.c_str());
        } else if (n == "device") {
            uint64_t d = 0;
            if (!parse_u64(val, d))
                return invalid();
            a.device = (int)d;
        }
    }
    return 0;
}

size_t istream_read(void *ctx, void *dst, size_t n)
{
    auto *is = static_cast<std::istream *>(ctx);
    is->read(static_cast<char *>(dst), (std::streamsize)n);
    return (size_t)is->gcount();
}

bool read_all(std::istream &is, std::string &out)
{
    out.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
    return !is.bad();
}

// ---------------------------------------------------------------------------
// PNG (ISO/IEC 15948): IHDR, one IDAT, IEND.  Colour type 0 (gray) or 2
// (RGB), 8 or 16 bits per sample, no interlace.  Each row takes the filter
// (None, Sub, Up, Average, Paeth) whose output has the smallest sum of
// absolute values as signed bytes, the heuristic the specification suggests.
// ---------------------------------------------------------------------------

void put_u32(std::string &s, uint32_t v)
{
    s.push_back((char)(v >> 24));
    s.push_back((char)(v >> 16));
    s.push_back((char)(v >> 8));
    s.push_back((char)v);
}

void put_chunk(std::string &png, const char *type, const unsigned char *data, size_t n)
{
    put_u32(png, (uint32_t)n);
    const size_t start = png.size();
    png.append(type, 4);
    if (n)
        png.append((const char *)data, n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32_z(crc, (const Bytef *)png.data() + start, (z_size_t)(4 + n));
    put_u32(png, (uint32_t)crc);
}

inline unsigned char paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc)
        return (unsigned char)a;
    return (unsigned char)(pb <= pc ? b : c);
}

// pixels: host-order samples, row-major; returns false when zlib fails
bool encode_png(const void *pixels, uint32_t w, uint32_t h, uint32_t channels,
                uint32_t bytes_per_sample, std::string &png)
{
    const size_t bpp = (size_t)channels * bytes_per_sample; // bytes per pixel
    const size_t stride = (size_t)w * bpp;
    // big-endian sample bytes, row by row
    std::vector<unsigned char> raw(stride * h);
    if (bytes_per_sample == 1) {
        std::memcpy(raw.data(), pixels, raw.size());
    } else {
        const uint16_t *p = (const uint16_t *)pixels;
        for (size_t i = 0; i < raw.size() / 2; ++i) {
            raw[2 * i] = (unsigned char)(p[i] >> 8);
            raw[2 * i + 1] = (unsigned char)p[i];
        }
    }
    std::vector<unsigned char> filt((stride + 1) * h);
    std::vector<unsigned char> cand[5];
    for (auto &c : cand)
        c.resize(stride);
    std::vector<unsigned char> zero(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const unsigned char *cur = raw.data() + (size_t)y * stride;
        const unsigned char *up = y ? cur - stride : zero.data();
        uint64_t best_sum = ~0ull;
        int best = 0;
        for (int f = 0; f < 5; ++f) {
            unsigned char *o = cand[f].data();
            uint64_t sum = 0;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0;
                const int b = up[i];
                const int c = i >= bpp ? up[i - bpp] : 0;
                int pred = 0;
                switch (f) {
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: break;
                }
                o[i] = (unsigned char)(cur[i] - pred);
                sum += (uint64_t)std::abs((int)(signed char)o[i]);
            }
            if (sum < best_sum) {
                best_sum = sum;
                best = f;
            }
        }
        unsigned char *row = filt.data() + (size_t)y * (stride + 1);
        row[0] = (unsigned char)best;
        std::memcpy(row + 1, cand[best].data(), stride);
    }
    uLongf zlen = compressBound((uLong)filt.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, filt.data(), (uLong)filt.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
        return false;

    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    png.assign((const char *)sig, 8);
    unsigned char ihdr[13];
    const uint32_t dims[2] = {w, h};
    for (int k = 0; k < 2; ++k)
        for (int b = 0; b < 4; ++b)
            ihdr[4 * k + b] = (unsigned char)(dims[k] >> (24 - 8 * b));
    ihdr[8] = (unsigned char)(8 * bytes_per_sample); // bit depth
    ihdr[9] = channels == 3 ? 2 : 0;                 // colour type
    ihdr[10] = 0;                                    // deflate
    ihdr[11] = 0;                                    // adaptive filtering
    ihdr[12] = 0;                                    // no interlace
    put_chunk(png, "IHDR", ihdr, sizeof ihdr);
    // a chunk holds at most 2^31 - 1 bytes: large images take several
    // IDAT chunks, whose data concatenate into the one zlib stream
    const size_t max_chunk = (size_t)1 << 30;
    for (size_t off = 0; off < (size_t)zlen; off += max_chunk)
        put_chunk(png, "IDAT", z.data() + off, std::min(max_chunk, (size_t)zlen - off));
    put_chunk(png, "IEND", nullptr, 0);
    return true;
}

} // namespace

int main(int argc, char **argv)
{
    Args a;
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
    // ffr_img.cpp:92-97 (the reference throws; here: the message, exit 1).
    // The float build reads gamma as a float.
    const bool f32 = a.word == 32;
    if (f32 && !a.gamma_text.empty())
        a.gamma = (double)std::strtof(a.gamma_text.c_str(), nullptr);
    if (f32 ? (float)a.gamma < 1e-10f : a.gamma < 1e-20) {
        std::cerr << "ERROR: gamma too small" << std::endl;
        return 1;
    }
    if (a.bits != 8 && a.bits != 16) {
        std::cerr << "ERROR: bits per channel must be 8 or 16" << std::endl;
        return 1;
    }

    // echo the options (ffr_img.cpp:100-114)
    std::cerr << "ffr-img version " << kVersion << std::endl;
    std::cerr << "--flame " << a.flame << std::endl;
    for (auto &s : a.inputs)
        std::cerr << "--input " << s << std::endl;
    std::cerr << "--output " << a.output << std::endl;
    std::cerr << "--gamma " << a.gamma << std::endl;
    if (a.mono)
        std::cerr << "--monochrome" << std::endl;
    if (a.gray)
        std::cerr << "--grayscale" << std::endl;
    if (a.color)
        std::cerr << "--color" << std::endl;
    std::cerr << "--bits " << a.bits << std::endl;
    std::cerr << "--" << std::endl;

    // the flame (ffr_img.cpp:116-135)
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
        std::cerr << "flame: " << doc.dump() << std::endl;
        dims = doc["dimensions"].intValue();
    } catch (const std::exception &e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 1;
    }
    if (dims != 2) {
        std::cerr << "ERROR: only 2D flames supported" << std::endl;
        return 1;
    }

    // the renderer holds the summed buffer (ffr_img.cpp:137-149)
    ffr_options opt;
    ffr_default_options(&opt);
    opt.word = f32 ? FFR_WORD32 : FFR_WORD64;
    opt.precision = f32 ? FFR_FP32 : FFR_FP64;
    opt.device = a.device;
    ffr_renderer *r = nullptr;
    if (ffr_create(text.data(), text.size(), &opt, &r) != FFR_OK) {
        std::cerr << "ERROR: " << ffr_last_error() << std::endl;
        return 1;
    }
    const size_t bytes = ffr_buffer_bytes(r);
    std::cerr << "buffer: " << bytes << " bytes, little endian, " << ffr_word_bytes(r)
              << " byte numbers" << std::endl;
    std::cerr << "color: " << ffr_color_dims(r) << " dimensions" << std::endl;

    // ffr_img.cpp:151-165
    if ((int)a.mono + (int)a.gray + (int)a.color != 1) {
        std::cerr << "ERROR: must choose exactly 1 coloring flag "
                  << "(--monochrome, --grayscale, --color)" << std::endl;
        ffr_destroy(r);
        return 1;
    }
    if (a.inputs.empty()) {
        std::cerr << "ERROR: no input buffer files specified" << std::endl;
        ffr_destroy(r);
        return 1;
    }

    // add the input buffers (ffr_img.cpp:166-183), streamed to the device
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

    // render_image (ffr_img.cpp:199-309): bounds, then the pixels
    std::cerr << "processing buffer" << std::endl;
    const int mode = a.mono ? FFR_IMG_MONO : (a.gray ? FFR_IMG_GRAY : FFR_IMG_COLOR);
    const int bits = (int)a.bits;
    ffr_image_info info;
    std::vector<unsigned char> pix(ffr_image_bytes(r, mode, bits));
    int rc = ffr_render_image(r, mode, a.gamma, bits, nullptr, 0, &info);
    if (rc == FFR_OK || rc == FFR_E_EMPTY) {
        std::cerr << "histogram bounds: " << info.hist_min << " " << info.hist_max << std::endl;
        std::cerr << "scaler bounds: " << info.scale_min << " " << info.scale_max << std::endl;
    }
    if (rc == FFR_OK)
        rc = ffr_render_image(r, mode, a.gamma, bits, pix.data(), pix.size(), &info);
    if (rc != FFR_OK) {
        std::cerr << "ERROR: " << ffr_last_error() << std::endl;
        ffr_destroy(r);
        return 1;
    }
    ffr_destroy(r);

    std::string png;
    if (!encode_png(pix.data(), info.width, info.height, info.channels, info.bytes_per_channel,
                    png)) {
        std::cerr << "ERROR: error writing output" << std::endl;
        return 1;
    }
    bool ok;
    if (a.output == "-") {
        std::cout.write(png.data(), (std::streamsize)png.size());
        std::cout.flush();
        ok = std::cout.good();
    } else {
        std::ofstream f(a.output, std::ios::out | std::ios::binary);
        f.write(png.data(), (std::streamsize)png.size());
        ok = f.good();
    }
    if (!ok) {
        std::cerr << "ERROR: error writing output" << std::endl;
        return 1;
    }
    return 0;
}

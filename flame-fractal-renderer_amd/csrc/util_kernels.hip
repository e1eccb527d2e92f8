// Buffer and statistics kernels used by the C ABI.
#include "util_kernels.hpp"

#include <algorithm>

namespace ffr {

__global__ void reset_stats_kernel(DevStats *st, unsigned long long *xf_count, int nxf)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        st->s_iter = 0;
        st->s_plot = 0;
        st->n_bad = 0;
        st->atomics = 0;
        st->flush_atomics = 0;
        for (int i = 0; i < 3; ++i) {
            st->pmin[i] = (double)INFINITY;
            st->pmax[i] = -(double)INFINITY;
        }
    }
    for (int k = t; k < nxf; k += gridDim.x * blockDim.x)
        xf_count[k] = 0;
}

// the statistics of several renderers and the task-queue counters of a
// batched launch: block b clears entry b, block 0 also the counters
__global__ void reset_many_kernel(const ResetEntry *e, int n, unsigned *queues, int nq)
{
    const int b = blockIdx.x;
    if (b < n) {
        DevStats *st = e[b].st;
        if (threadIdx.x == 0) {
            st->s_iter = 0;
            st->s_plot = 0;
            st->n_bad = 0;
            st->atomics = 0;
        st->flush_atomics = 0;
            for (int i = 0; i < 3; ++i) {
                st->pmin[i] = (double)INFINITY;
                st->pmax[i] = -(double)INFINITY;
            }
        }
        for (int k = threadIdx.x; k < e[b].nxf; k += blockDim.x)
            e[b].xf[k] = 0;
    }
    if (b == 0)
        for (int i = threadIdx.x; i < nq; i += blockDim.x)
            queues[i] = 0;
}

// addBuffer (buffer_renderer.hpp:375-392): word 0 of a cell adds as an integer,
// the colour words add as floating point
template <typename CT, typename CR>
__global__ void add_cells_kernel(CT *dst, const CT *src, unsigned long long words, int cw)
{
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         i < words; i += (unsigned long long)gridDim.x * blockDim.x) {
        if (i % (unsigned long long)cw == 0) {
            dst[i] += src[i];
        } else {
            CR *d = reinterpret_cast<CR *>(dst + i);
            const CR *s = reinterpret_cast<const CR *>(src + i);
            *d = *d + *s;
        }
    }
}

// reductions run 256-lane workgroups (four waves) and finish with one
// atomic per workgroup
constexpr int kReduceBlock = 256;
constexpr int kReduceWaves = kReduceBlock / 64;
constexpr int kReduceGrid = 512;

template <typename CT>
__global__ void hist_stats_kernel(const CT *buf, unsigned long long cells, int cw,
                                  unsigned long long *out /* sum, min, max */)
{
    unsigned long long sum = 0, mn = ~0ull, mx = 0;
#pragma unroll 4
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         c < cells; c += (unsigned long long)gridDim.x * blockDim.x) {
        unsigned long long v = (unsigned long long)buf[c * (unsigned long long)cw];
        sum += v;
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
        unsigned long long a = __shfl_xor(mn, off, 64);
        unsigned long long b = __shfl_xor(mx, off, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    // one set of atomics per workgroup: atomics on one address serialise
    // (about 80 M/s), so per-wave atomics over a 4K buffer took milliseconds
    __shared__ unsigned long long ps[kReduceWaves], pn[kReduceWaves], px[kReduceWaves];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ps[w] = sum;
        pn[w] = mn;
        px[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            sum += ps[i];
            mn = pn[i] < mn ? pn[i] : mn;
            mx = px[i] > mx ? px[i] : mx;
        }
        atomicAdd(&out[0], sum);
        atomicMin(&out[1], mn);
        atomicMax(&out[2], mx);
    }
}

// L1 distance of two normalised count histograms with the same cells:
// sum over c of |a_c * inv_a - b_c * inv_b|, in double
template <typename CA, typename CB>
__global__ void l1_counts_kernel(const CA *a, int cwa, const CB *b, int cwb,
                                 unsigned long long cells, double inv_a, double inv_b,
                                 double *out)
{
    double s = 0.0;
#pragma unroll 4
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         c < cells; c += (unsigned long long)gridDim.x * blockDim.x) {
        const double d = (double)a[c * (unsigned long long)cwa] * inv_a -
                         (double)b[c * (unsigned long long)cwb] * inv_b;
        s += d < 0.0 ? -d : d;
    }
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, 64);
    __shared__ double part[kReduceWaves];
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            s += part[i];
        atomicAdd(out, s);
    }
}

__global__ void zero_double_kernel(double *p) { *p = 0.0; }

// The image stage of ffr-img (ffr_img.cpp:199-309, image_renderer.hpp:
// 150-192): one pixel per cell, rows in buffer order, in num_t arithmetic
// (N): mono  count != 0 -> 1;  gray  pow(log(1 + count) / max, gp);
// colour  that value times each colour sum / count.  A value v becomes the
// pixel (P)(v * scale), scale = 2^bits (1 - 2^-mantissa).  max = log(1 +
// largest count) is taken here, by the same log as every cell's, so the
// brightest cell maps to exactly 1.

// (pix_t)v as gcc compiles it for x86-64: a signed 32-bit conversion
// (cvttsd2si / cvttss2si) of which the pixel keeps the low bits; NaN and
// values outside the int32 range give INT32_MIN there, where the device's
// own conversion would clamp
template <typename N>
__device__ inline int32_t host_i32(N v)
{
    return (v >= (N)-2147483648.0 && v < (N)2147483648.0) ? (int32_t)v : INT32_MIN;
}

// glibc's logf and powf (the float build's std::log / std::pow) evaluate
// in double and round once, so they return the correctly rounded float in
// all but a vanishing share of arguments; the device does the same
__device__ inline double img_log(double x) { return log(x); }
__device__ inline float img_log(float x) { return (float)log((double)x); }
__device__ inline double img_pow(double x, double y) { return pow(x, y); }
__device__ inline float img_pow(float x, float y) { return (float)pow((double)x, (double)y); }

template <typename CT, typename CR, typename N, typename P>
__global__ void tonemap_kernel(const CT *buf, unsigned long long cells, int cw, int mode,
                               N maxc, N gp, N scale, P *out)
{
    const N max = img_log((N)1 + maxc);
    // an empty cell is 0 in every mode while gp > 0: log(1) / max = 0 and
    // pow(0, gp) = 0 in gray; in colour the sums over a zero count give
    // NaN or infinity, whose conversion leaves 0
    const bool empty_is_zero = gp > (N)0;
    for (unsigned long long c = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         c < cells; c += (unsigned long long)gridDim.x * blockDim.x) {
        const CT *cell = buf + c * (unsigned long long)cw;
        const CT h = cell[0];
        if (mode == 0) {
            out[c] = (P)host_i32((h != 0 ? (N)1.0 : (N)0.0) * scale);
            continue;
        }
        if (mode == 1) {
            N v = (N)0;
            if (h != 0 || !empty_is_zero)
                v = img_pow(img_log((N)1 + (N)h) / max, gp);
            out[c] = (P)host_i32(v * scale);
            continue;
        }
        P px[3] = {0, 0, 0};
        if (h != 0) {
            const N ll = img_pow(img_log((N)1 + (N)h) / max, gp);
            const N hn = (N)h;
            const CR *col = reinterpret_cast<const CR *>(cell + 1);
            for (int i = 0; i < 3; ++i)
                px[i] = (P)host_i32(ll * ((N)col[i] / hn) * scale);
        }
        for (int i = 0; i < 3; ++i)
            out[c * 3 + i] = px[i];
    }
}

// Cell updates at cells chosen by a hash of the update index (a 64-bit
// finaliser, uniform enough over any cell count), each one integer add on
// the count word and a float add on every colour word, with the walker
// kernel's instructions and scope.
template <typename CT, typename CR>
__global__ void scatter_atomics_kernel(CT *buf, unsigned long long cells, int cw,
                                       unsigned long long updates)
{
    for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         u < updates; u += (unsigned long long)gridDim.x * blockDim.x) {
        unsigned long long h = u * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        const unsigned long long c = __umul64hi(h, cells); // in [0, cells)
        CT *cell = buf + c * (unsigned long long)cw;
        __hip_atomic_fetch_add(cell, (CT)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        CR *col = reinterpret_cast<CR *>(cell + 1);
        for (int i = 0; i < cw - 1; ++i)
            unsafeAtomicAdd(&col[i], (CR)0.5);
    }
}

__global__ void hist_stats_init_kernel(unsigned long long *out)
{
    out[0] = 0;
    out[1] = ~0ull;
    out[2] = 0;
}

__global__ void fill_zero_kernel(unsigned int *p, unsigned long long n)
{
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

static int grid_for(unsigned long long n, int block)
{
    unsigned long long g = (n + (unsigned long long)block - 1) / (unsigned long long)block;
    if (g > 65536ull)
        g = 65536ull;
    return g ? (int)g : 1;
}

int launch_reset_stats(DevStats *st, unsigned long long *xf_count, int nxf, hipStream_t s)
{
    hipLaunchKernelGGL(reset_stats_kernel, dim3(1), dim3(256), 0, s, st, xf_count, nxf);
    return (int)hipGetLastError();
}

int launch_reset_many(const ResetEntry *e, int n, unsigned *queues, int nq, hipStream_t s)
{
    hipLaunchKernelGGL(reset_many_kernel, dim3(n > 0 ? n : 1), dim3(256), 0, s, e, n, queues,
                       nq);
    return (int)hipGetLastError();
}

int launch_add_cells(void *dst, const void *src, unsigned long long words, int cw,
                     int word_bytes, hipStream_t s)
{
    const int block = 256;
    const int grid = grid_for(words, block);
    if (word_bytes == 8)
        hipLaunchKernelGGL((add_cells_kernel<unsigned long long, double>), dim3(grid),
                           dim3(block), 0, s, (unsigned long long *)dst,
                           (const unsigned long long *)src, words, cw);
    else
        hipLaunchKernelGGL((add_cells_kernel<unsigned int, float>), dim3(grid), dim3(block), 0,
                           s, (unsigned int *)dst, (const unsigned int *)src, words, cw);
    return (int)hipGetLastError();
}

int launch_hist_stats(const void *buf, unsigned long long cells, int cw, int word_bytes,
                      unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(hist_stats_init_kernel, dim3(1), dim3(1), 0, s, out);
    const int block = kReduceBlock;
    const int grid = std::min(grid_for(cells, block), kReduceGrid);
    if (word_bytes == 8)
        hipLaunchKernelGGL((hist_stats_kernel<unsigned long long>), dim3(grid), dim3(block), 0,
                           s, (const unsigned long long *)buf, cells, cw, out);
    else
        hipLaunchKernelGGL((hist_stats_kernel<unsigned int>), dim3(grid), dim3(block), 0, s,
                           (const unsigned int *)buf, cells, cw, out);
    return (int)hipGetLastError();
}

int launch_l1_counts(const void *a, int cwa, int wba, const void *b, int cwb, int wbb,
                     unsigned long long cells, double inv_a, double inv_b, double *out,
                     hipStream_t s)
{
    const int block = kReduceBlock;
    const int grid = std::min(grid_for(cells, block), kReduceGrid);
    hipLaunchKernelGGL(zero_double_kernel, dim3(1), dim3(1), 0, s, out);
    using U8 = unsigned long long;
    using U4 = unsigned int;
    if (wba == 8 && wbb == 8)
        hipLaunchKernelGGL((l1_counts_kernel<U8, U8>), dim3(grid), dim3(block), 0, s,
                           (const U8 *)a, cwa, (const U8 *)b, cwb, cells, inv_a, inv_b, out);
    else if (wba == 8)
        hipLaunchKernelGGL((l1_counts_kernel<U8, U4>), dim3(grid), dim3(block), 0, s,
                           (const U8 *)a, cwa, (const U4 *)b, cwb, cells, inv_a, inv_b, out);
    else if (wbb == 8)
        hipLaunchKernelGGL((l1_counts_kernel<U4, U8>), dim3(grid), dim3(block), 0, s,
                           (const U4 *)a, cwa, (const U8 *)b, cwb, cells, inv_a, inv_b, out);
    else
        hipLaunchKernelGGL((l1_counts_kernel<U4, U4>), dim3(grid), dim3(block), 0, s,
                           (const U4 *)a, cwa, (const U4 *)b, cwb, cells, inv_a, inv_b, out);
    return (int)hipGetLastError();
}

int launch_tonemap(const void *buf, unsigned long long cells, int cw, int word_bytes, int mode,
                   unsigned long long maxc, double gp, int bits, void *out, hipStream_t s)
{
    // 2048 workgroups of 256 (eight waves per SIMD); each lane takes about
    // 16 cells of a 3840 x 2160 buffer, so the max above is taken rarely
    const int block = 256;
    const int grid = std::min(grid_for(cells, block), 2048);
    using U8 = unsigned long long;
    using U4 = unsigned int;
    if (word_bytes == 8) {
        // the default build: num_t = double, 64-bit words
        const double scale = (bits == 8 ? 256.0 : 65536.0) * (1.0 - 1.0 / 4503599627370496.0);
        if (bits == 8)
            hipLaunchKernelGGL((tonemap_kernel<U8, double, double, unsigned char>), dim3(grid),
                               dim3(block), 0, s, (const U8 *)buf, cells, cw, mode, (double)maxc, gp,
                               scale, (unsigned char *)out);
        else
            hipLaunchKernelGGL((tonemap_kernel<U8, double, double, unsigned short>), dim3(grid),
                               dim3(block), 0, s, (const U8 *)buf, cells, cw, mode, (double)maxc, gp,
                               scale, (unsigned short *)out);
    } else {
        // the float build: num_t = float, 32-bit words
        const float scale = (bits == 8 ? 256.0f : 65536.0f) * (1.0f - 1.0f / 8388608.0f);
        if (bits == 8)
            hipLaunchKernelGGL((tonemap_kernel<U4, float, float, unsigned char>), dim3(grid),
                               dim3(block), 0, s, (const U4 *)buf, cells, cw, mode, (float)maxc,
                               (float)gp, scale, (unsigned char *)out);
        else
            hipLaunchKernelGGL((tonemap_kernel<U4, float, float, unsigned short>), dim3(grid),
                               dim3(block), 0, s, (const U4 *)buf, cells, cw, mode, (float)maxc,
                               (float)gp, scale, (unsigned short *)out);
    }
    return (int)hipGetLastError();
}

int launch_scatter_atomics(void *buf, unsigned long long cells, int cw, int word_bytes,
                           unsigned long long updates, hipStream_t s)
{
    const int block = 256;
    const int grid = std::min(grid_for(updates, block), 8192);
    if (word_bytes == 8)
        hipLaunchKernelGGL((scatter_atomics_kernel<unsigned long long, double>), dim3(grid),
                           dim3(block), 0, s, (unsigned long long *)buf, cells, cw, updates);
    else
        hipLaunchKernelGGL((scatter_atomics_kernel<unsigned int, float>), dim3(grid), dim3(block),
                           0, s, (unsigned int *)buf, cells, cw, updates);
    return (int)hipGetLastError();
}

int launch_fill_zero(void *p, unsigned long long bytes, hipStream_t s)
{
    // buffers are whole words of 4 or 8 bytes
    const unsigned long long n = bytes / 4;
    const int block = 256;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(n, block)), dim3(block), 0, s,
                       (unsigned int *)p, n);
    return (int)hipGetLastError();
}

} // namespace ffr

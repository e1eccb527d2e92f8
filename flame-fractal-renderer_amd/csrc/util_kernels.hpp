// Launchers of the buffer and statistics kernels (util_kernels.hip).
#pragma once

#include "walk.hpp"

namespace ffr {

int launch_reset_stats(DevStats *st, unsigned long long *xf_count, int nxf, hipStream_t s);
int launch_add_cells(void *dst, const void *src, unsigned long long words, int cw,
                     int word_bytes, hipStream_t s);
int launch_hist_stats(const void *buf, unsigned long long cells, int cw, int word_bytes,
                      unsigned long long *out, hipStream_t s);
// zero a buffer with a kernel (usable inside stream capture)
int launch_fill_zero(void *p, unsigned long long bytes, hipStream_t s);

#define FFR_DECL_WALK(mode, R)                                                          \
    int launch_walk_##mode##_d1(int nc, int vs, bool cache, const WalkArgs<R> &a, int grid, \
                                int block, size_t lds, hipStream_t s);                    \
    int launch_walk_##mode##_d2(int nc, int vs, bool cache, const WalkArgs<R> &a, int grid, \
                                int block, size_t lds, hipStream_t s);                    \
    int launch_walk_##mode##_d3(int nc, int vs, bool cache, const WalkArgs<R> &a, int grid, \
                                int block, size_t lds, hipStream_t s);
FFR_DECL_WALK(f64u64, double)
FFR_DECL_WALK(f32u32, float)
FFR_DECL_WALK(f32u64, float)
#undef FFR_DECL_WALK

} // namespace ffr

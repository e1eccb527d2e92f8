// Walker kernels for 2-D flames, num_t = double, hist_t = uint64 (the reference's default build).
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f64u64_d2(int nc, int vs, bool cache, const WalkArgs<double> &a, int grid, int block,
                          size_t lds, hipStream_t s)
{
    return launch_nc<double, unsigned long long, double, 2>(nc, vs, cache, a, grid, block, lds, s);
}
} // namespace ffr

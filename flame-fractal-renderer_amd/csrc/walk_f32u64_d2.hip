// Walker kernels for 2-D flames, float trajectories, uint64 counts and double colour sums.
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f32u64_d2(int nc, int vs, bool cache, const WalkArgs<float> &a, int grid, int block,
                          size_t lds, hipStream_t s)
{
    return launch_nc<float, unsigned long long, double, 2>(nc, vs, cache, a, grid, block, lds, s);
}
} // namespace ffr

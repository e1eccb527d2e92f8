// Walker kernels for 3-D flames, num_t = float, hist_t = uint32 (the reference's float build).
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f32u32_d3(int nc, int vs, bool cache, const WalkArgs<float> &a, int grid, int block,
                          size_t lds, hipStream_t s)
{
    return launch_nc<float, unsigned int, float, 3>(nc, vs, cache, a, grid, block, lds, s);
}
} // namespace ffr

// Walker kernels for 3-D flames, num_t = float, hist_t = uint32 (the reference's float build).
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f32u32_d3(const WalkLaunch &L)
{
    return launch_walk<float, unsigned int, float, 3>(L);
}
} // namespace ffr

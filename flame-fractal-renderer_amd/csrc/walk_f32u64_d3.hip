// Walker kernels for 3-D flames, float trajectories, uint64 counts and double colour sums.
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f32u64_d3(const WalkLaunch &L)
{
    return launch_walk<float, unsigned long long, double, 3>(L);
}
} // namespace ffr

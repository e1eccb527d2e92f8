// Walker kernels for 2-D flames, float trajectories, uint64 counts and double colour sums.
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f32u64_d2(const WalkLaunch &L)
{
    return launch_walk<float, unsigned long long, double, 2>(L);
}
} // namespace ffr

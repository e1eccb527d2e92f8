// Walker kernels for 3-D flames, num_t = double, hist_t = uint64 (the reference's default build).
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f64u64_d3(const WalkLaunch &L)
{
    return launch_walk<double, unsigned long long, double, 3>(L);
}
} // namespace ffr

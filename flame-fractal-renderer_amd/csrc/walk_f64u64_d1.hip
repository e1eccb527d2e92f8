// Walker kernels for 1-D flames, num_t = double, hist_t = uint64 (the reference's default build).
#include "walk_kernel.inc"

namespace ffr {
int launch_walk_f64u64_d1(const WalkLaunch &L)
{
    return launch_walk<double, unsigned long long, double, 1>(L);
}
} // namespace ffr

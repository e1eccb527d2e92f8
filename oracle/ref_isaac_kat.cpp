// Known-answer driver for the reference's own ISAAC generator.
// Test infrastructure: built by `make -C oracle ref` from
// /root/reference/src/rng/isaac.hpp (the header is used where it lies; no
// reference source is copied).  Prints, for each seed, the first words the
// reference's Isaac<word_t,4> returns after setSeed(seed).
#include "rng/isaac.hpp"

#include <cinttypes>
#include <cstdio>

using tkoz::flame::Isaac;

int main()
{
    const uint64_t seeds64[] = {0ull, 1ull, 0x5eed0001ull, 0xdeadbeefcafef00dull,
                                0xffffffffffffffffull};
    const uint32_t seeds32[] = {0u, 1u, 0x5eed0003u, 0xffffffffu};
    const int n = 40; // spans three 16-word generations
    std::printf("{\n \"isaac64_u64seed\": {\n");
    for (size_t k = 0; k < sizeof seeds64 / sizeof seeds64[0]; ++k) {
        Isaac<uint64_t, 4> r(seeds64[k]);
        std::printf("  \"%" PRIu64 "\": [", seeds64[k]);
        for (int i = 0; i < n; ++i)
            std::printf("%s\"%016" PRIx64 "\"", i ? "," : "", r.next());
        std::printf("]%s\n", k + 1 < sizeof seeds64 / sizeof seeds64[0] ? "," : "");
    }
    std::printf(" },\n \"isaac32_u32seed\": {\n");
    for (size_t k = 0; k < sizeof seeds32 / sizeof seeds32[0]; ++k) {
        Isaac<uint32_t, 4> r(seeds32[k]);
        std::printf("  \"%" PRIu32 "\": [", seeds32[k]);
        for (int i = 0; i < n; ++i)
            std::printf("%s\"%08" PRIx32 "\"", i ? "," : "", r.next());
        std::printf("]%s\n", k + 1 < sizeof seeds32 / sizeof seeds32[0] ? "," : "");
    }
    std::printf(" },\n \"isaac32_u64seed\": {\n");
    for (size_t k = 0; k < sizeof seeds64 / sizeof seeds64[0]; ++k) {
        Isaac<uint32_t, 4> r(seeds64[k]);
        std::printf("  \"%" PRIu64 "\": [", seeds64[k]);
        for (int i = 0; i < n; ++i)
            std::printf("%s\"%08" PRIx32 "\"", i ? "," : "", r.next());
        std::printf("]%s\n", k + 1 < sizeof seeds64 / sizeof seeds64[0] ? "," : "");
    }
    std::printf(" }\n}\n");
    return 0;
}

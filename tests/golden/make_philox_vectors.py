"""Write tests/golden/philox_vectors.json from rocRAND's Philox4x32-10.

rocRAND ships Philox4x32-10 (Salmon et al., SC'11) as a header-only engine
(/opt/rocm/include/rocrand/rocrand_philox4x32_10.h).  This script compiles a
small host program against that header with hipcc and records
(counter, key) -> output blocks, the library's answers for the walker-stream
generator.  The output is data; the test checks the oracle's generator and
the stream layout against it.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main() {
    Engine e;
    unsigned s = 12345u;
    auto next = [&]() { s = s * 1664525u + 1013904223u; return s; };
    std::printf("[");
    for (int i = 0; i < 64; ++i) {
        uint4 c = {next(), next(), next(), next()};
        if (i < 8) { c.y = 0; c.w = 0; }
        uint2 k = {next(), next()};
        uint4 o = e.ten_rounds(c, k);
        std::printf("%s[[%u,%u,%u,%u],[%u,%u],[%u,%u,%u,%u]]", i ? "," : "",
                    c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w);
    }
    std::printf("]\n");
    return 0;
}
"""


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-x", "hip", "--offload-arch=gfx950",
                               src, "-o", exe])
        rows = json.loads(subprocess.check_output([exe]))
    out = {
        "provenance": "rocRAND philox4x32_10_engine::ten_rounds (host), "
                      "ROCm 7.2 header rocrand_philox4x32_10.h; tests/golden/make_philox_vectors.py",
        "vectors": [{"ctr": c, "key": k, "out": o} for c, k, o in rows],
    }
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    return 0


if __name__ == "__main__":
    sys.exit(main())

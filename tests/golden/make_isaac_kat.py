"""Write tests/golden/isaac_kat.json from the reference's own ISAAC.

Runs oracle/_ref/isaac_kat, which `make -C oracle ref` builds from the
reference's src/rng/isaac.hpp (in this container only).  The output is data:
the first 40 words of Isaac<u64,4> and Isaac<u32,4> for a few seeds.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "isaac_kat")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    data = json.loads(subprocess.check_output([exe]))
    data["provenance"] = ("first 40 words of the reference's Isaac<word_t,4> after "
                          "setSeed(seed); src/rng/isaac.hpp compiled by "
                          "oracle/Makefile (target ref)")
    with open(os.path.join(HERE, "isaac_kat.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

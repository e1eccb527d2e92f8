"""Pins of the oracle's random numbers.

- ISAAC (the reference's generator, src/rng/isaac.hpp): bit-exact against the
  reference header itself (fixture written by tests/golden/make_isaac_kat.py).
- Philox4x32-10 (the walker streams): the Random123 published known answers
  and rocRAND's implementation (fixture from tests/golden/make_philox_vectors.py).
- The stream layout and randNum conversions: closed forms.
"""

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_isaac64_matches_reference(oracle_mod):
    kat = _load("isaac_kat.json")
    for seed, words in kat["isaac64_u64seed"].items():
        got = ["%016x" % w for w in oracle_mod.isaac64_words(int(seed), len(words))]
        assert got == words, seed


def test_isaac32_matches_reference(oracle_mod):
    kat = _load("isaac_kat.json")
    for seed, words in kat["isaac32_u32seed"].items():
        got = ["%08x" % w for w in oracle_mod.isaac32_words(int(seed), len(words), 32)]
        assert got == words, seed
    for seed, words in kat["isaac32_u64seed"].items():
        got = ["%08x" % w for w in oracle_mod.isaac32_words(int(seed), len(words), 64)]
        assert got == words, seed


# Random123 kat_vectors (D. E. Shaw Research), philox4x32 with 10 rounds
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123(oracle_mod, ctr, key, out):
    assert oracle_mod.philox4x32_10(ctr, key) == out


def test_philox_matches_rocrand(oracle_mod):
    vec = _load("philox_vectors.json")["vectors"]
    assert len(vec) >= 64
    for v in vec:
        assert oracle_mod.philox4x32_10(v["ctr"], v["key"]) == v["out"], v


def test_stream_layout(oracle_mod):
    """Word n of walker w: half (n & 1) of the block at counter
    (lo n>>1, hi n>>1, lo w, hi w) under key (lo seed, hi seed)."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        seed = int(rng.integers(0, 2 ** 63)) * 2 + 1
        w = int(rng.integers(0, 2 ** 62))
        n = int(rng.integers(0, 2 ** 40))
        blk = n >> 1
        o = oracle_mod.philox4x32_10(
            [blk & 0xFFFFFFFF, blk >> 32, w & 0xFFFFFFFF, w >> 32],
            [seed & 0xFFFFFFFF, seed >> 32])
        want = (o[0] | (o[1] << 32)) if n % 2 == 0 else (o[2] | (o[3] << 32))
        assert oracle_mod.stream_word(seed, w, n) == want


def test_stream_words_distinct_across_walkers(oracle_mod):
    a = [oracle_mod.stream_word(7, 0, n) for n in range(64)]
    b = [oracle_mod.stream_word(7, 1, n) for n in range(64)]
    assert len(set(a) | set(b)) == 128


def test_randnum_closed_form_and_first_point(oracle_mod):
    """A walker's start point is 2u - 1 with u = (w >> 11) / 2^53 (double)
    or (w >> 40) / 2^24 (float), from its first words (flame_rng.hpp:67-87,
    151-158): pinned by a one-xform flame with zero settle iterations whose
    map is the constant b = start, so the extremes report the start point."""
    # linear with A = I and no settle: p1 = p0 = 2u - 1 per coordinate
    text = json.dumps({"dimensions": 2, "size": [4, 4], "bounds": [[-1, 1], [-1, 1]],
                       "xforms": [{"weight": 1, "variations": [{"name": "linear", "weight": 1}]}]})
    for mode, conv in ((oracle_mod.F64_U64, lambda w: (w >> 11) / 2.0 ** 53),
                       (oracle_mod.F32_U64, lambda w: float(np.float32((w >> 40) / 2.0 ** 24)))):
        f = oracle_mod.OracleFlame(text, mode, settle=0)
        for walker in range(8):
            r = f.render(1, 1, seed=99, walker_base=walker)
            # draw 0, 1: start point; draw 2: xform selection
            want = [2.0 * conv(oracle_mod.stream_word(99, walker, i)) - 1.0 for i in range(2)]
            if mode != oracle_mod.F64_U64:
                want = [float(np.float32(v)) for v in want]
            assert list(r.pt_min) == want and list(r.pt_max) == want

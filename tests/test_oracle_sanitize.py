"""The CPU oracle under sanitizers (CPU only).

Every GPU parity test trusts the oracle, and several run its thread pool:
walkers claimed under a mutex, relaxed atomic adds on the shared buffer, CAS
loops for colour words, a mutex around the bad-value records (the
restatement of BufferRenderer::render, src/renderers/buffer_renderer.hpp:
269-338).  `make -C oracle sanitize` builds oracle/sanitize_check.c with the
oracle compiled in, once under AddressSanitizer + UndefinedBehaviorSanitizer
and once under ThreadSanitizer.  The driver renders a linear flame, a colour
flame of transcendental and word-drawing variations with a final xform, and
a flame with bad values, in all three precision modes, on six threads and on
one, and requires the two to agree.  Any sanitizer report fails the run.
"""

import os
import subprocess

import pytest

from conftest import ROOT

ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def built():
    try:
        subprocess.run(["make", "-C", ORACLE, "sanitize"], check=True, capture_output=True,
                       timeout=300)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"sanitizer build unavailable: {e}")
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("kind,env", [
    ("asan", {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1:exitcode=66",
              "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}),
    ("tsan", {"TSAN_OPTIONS": "halt_on_error=1:exitcode=66"}),
])
def test_oracle_pool_is_clean_under_sanitizers(built, kind, env):
    e = dict(os.environ)
    e.update(env)
    p = subprocess.run([os.path.join(built, f"check_{kind}")], capture_output=True, text=True,
                       timeout=300, env=e)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "sanitize check ok" in p.stdout
    assert "Sanitizer" not in out, out[-4000:]
    # the bad-value flame exercised the locked records in every mode
    bad_lines = [ln for ln in p.stdout.splitlines() if ln.startswith("occasional_bad")]
    assert len(bad_lines) == 3
    for ln in bad_lines:
        assert int(ln.split(" bad,")[0].split()[-1]) > 100, ln

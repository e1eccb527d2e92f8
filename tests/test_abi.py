"""The C ABI library and the ffr-buf program, without a GPU.

Checks that libffr_hip.so loads and exports every function include/ffr.h
declares, that flame validation works without a device, and the CLI's
argument and flame error paths (which run before any device call).
"""

import json
import os
import re
import subprocess

import pytest

import flames as TF
from conftest import PKG, ROOT


def declared_functions():
    with open(os.path.join(ROOT, "include", "ffr.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function(ffr_mod):
    names = declared_functions()
    assert len(names) >= 25
    lib = ffr_mod.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(ffr_mod.EXPORTS) == names


def test_default_options(ffr_mod):
    import ctypes as C
    o = ffr_mod.Options()
    ffr_mod.lib().ffr_default_options(C.byref(o))
    assert (o.precision, o.word, o.settle_iters, o.bad_value_limit) == (0, 0, -1, 256)
    assert (o.cell_cache, o.cache_kib, o.block, o.binned, o.compiled) == (-1, 0, 0, -1, -1)
    # the header's struct: 4 x 32-bit, 4 x 64-bit, 6 x 32-bit fields
    assert C.sizeof(o) == 72


def test_check_flame_without_device(ffr_mod):
    ids, cw = ffr_mod.check_flame(TF.fern())
    assert ids == [1, 2, 3, 0]
    assert cw[-1] == 1.0


def test_no_torch_or_hip_types_in_header():
    with open(os.path.join(ROOT, "include", "ffr.h")) as f:
        text = f.read()
    for banned in ("hipStream_t", "torch", "at::", "std::"):
        assert banned not in re.sub(r"/\*.*?\*/", "", text, flags=re.S), banned


def _cli(args, stdin=None):
    exe = os.path.join(PKG, "ffr-buf")
    assert os.path.exists(exe), "ffr-buf is not built"
    p = subprocess.run([exe] + args, input=stdin, capture_output=True, timeout=60)
    return p.returncode, p.stdout, p.stderr.decode(errors="replace")


def test_cli_help_and_missing_options():
    rc, out, err = _cli([])
    assert rc == 1 and "--flame" in err and out == b""
    rc, _, err = _cli(["-h"])
    assert rc == 1 and "--bad-values" in err
    rc, _, err = _cli(["-o", "-"])
    assert rc != 0 and "'--flame' is required" in err


def test_cli_flame_errors(tmp_path):
    f = tmp_path / "f.json"
    f.write_text(json.dumps({"dimensions": 5}))
    rc, _, err = _cli(["-f", str(f), "-o", "-"])
    assert rc == 1 and "ERROR: 5D not supported" in err
    assert "--samples 0 (not rendering)" in err and "flame: {\"dimensions\":5}" in err
    f.write_text(TF.fern().replace('"linear"', '"nope"', 1))
    rc, _, err = _cli(["-f", str(f), "-o", "-"])
    assert rc != 0 and "unknown variation: nope" in err
    f.write_text("{ not json")
    rc, _, err = _cli(["-f", str(f), "-o", "-"])
    assert rc != 0 and "parse error" in err


def test_cli_flame_echo_format(tmp_path):
    f = tmp_path / "f.json"
    f.write_text('{"dimensions": 4, "b": [1e-20, 100000.0, 0.5, 1e+21, -0.0, 3], "a": "x"}')
    rc, _, err = _cli(["-f", str(f), "-o", "-"])
    # nlohmann-style dump: sorted keys, shortest doubles, ".0" on integral values
    assert 'flame: {"a":"x","b":[1e-20,100000.0,0.5,1e+21,-0.0,3],"dimensions":4}' in err


def _img(args, stdin=None):
    exe = os.path.join(PKG, "ffr-img")
    assert os.path.exists(exe), "ffr-img is not built"
    p = subprocess.run([exe] + args, input=stdin, capture_output=True, timeout=60)
    return p.returncode, p.stdout, p.stderr.decode(errors="replace")


def test_img_cli_help_and_option_errors(tmp_path):
    """ffr-img's checks before any device work (ffr_img.cpp:80-135)"""
    rc, out, err = _img([])
    assert rc == 1 and "--grayscale" in err and out == b""
    rc, _, err = _img(["-h"])
    assert rc == 1 and "--gamma" in err
    rc, _, err = _img(["-f", "x.json"])
    assert rc == 1 and "'--output' is required" in err
    rc, _, err = _img(["-f", "x.json", "-o", "-", "-y", "0"])
    assert rc == 1 and "gamma too small" in err
    rc, _, err = _img(["-f", "x.json", "-o", "-", "-y", "1e-21"])
    assert rc == 1 and "gamma too small" in err
    rc, _, err = _img(["-f", "x.json", "-o", "-", "-b", "12"])
    assert rc == 1 and "bits per channel must be 8 or 16" in err
    rc, _, err = _img(["-f", "x.json", "-o", "-", "-m", "-b16", "-y", "2.2"])
    assert "--gamma 2.2\n--monochrome\n--bits 16\n--\n" in err
    f = tmp_path / "f.json"
    f.write_text(TF.tetrahedron())
    rc, _, err = _img(["-f", str(f), "-o", "-", "-g", "-i", "a.buf"])
    assert rc == 1 and "ERROR: only 2D flames supported" in err
    assert "--input a.buf\n" in err and "flame: {" in err

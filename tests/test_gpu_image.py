"""GPU parity of the image stage (ffr-img) against the CPU oracle.

The device tonemaps the buffer it holds; the oracle tonemaps a host copy of
the same buffer.  Bars:
  - mono: bit-exact;
  - gray and colour: the device's log / pow and glibc's may differ in the
    last bit, which moves a pixel only when its exact value lies within an
    ulp of a boundary; the bar is |difference| <= 1 on at most 1e-4 of the
    channels, and in practice every pixel is equal;
  - the bounds ffr-img prints: exact (counts) and equal to 1e-15 relative
    (log(1 + count) in double; 1e-6 for the float build);
  - the PNG that ffr-img writes decodes (by a decoder written here from the
    PNG specification, and by Pillow where it reads the format) to the pixels
    of the API.
"""

import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import flames as TF
from conftest import PKG, read_flame

pytestmark = pytest.mark.gpu

SEED = 0x1A6E


def render(ffr, text, samples, word=64, batch=1024, seed=SEED):
    precision = "fp64" if word == 64 else "fp32"
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=seed, batch_size=batch)
    assert r.render(samples)
    return r


def compare(gpu, ref, exact=False, label=""):
    assert gpu.shape == ref.shape and gpu.dtype == ref.dtype
    d = np.abs(gpu.astype(np.int64) - ref.astype(np.int64))
    where = [(tuple(int(i) for i in ix), int(gpu[ix]), int(ref[ix]))
             for ix in zip(*np.nonzero(d))][:8]
    if exact:
        assert int(d.max()) == 0, (label, where)
    else:
        assert int(d.max()) <= 1, (label, where)
        assert np.count_nonzero(d) <= max(1, d.size // 10000), (label, where)


def decode_png(data: bytes) -> np.ndarray:
    """PNG (colour types 0 / 2, depth 8 / 16, no interlace) to an array, per
    the specification: chunk CRCs, zlib stream, the five row filters"""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) & 0xFFFFFFFF == crc, kind
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, comp, filt, inter = ihdr
    assert comp == 0 and filt == 0 and inter == 0 and ctype in (0, 2) and depth in (8, 16)
    ch = 3 if ctype == 2 else 1
    bpp = ch * depth // 8
    stride = w * bpp
    raw = zlib.decompress(idat)
    assert len(raw) == h * (stride + 1)
    out = np.zeros((h, stride), dtype=np.int64)
    prev = np.zeros(stride, dtype=np.int64)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, dtype=np.uint8, count=stride,
                             offset=y * (stride + 1) + 1).astype(np.int64)
        cur = np.zeros(stride, dtype=np.int64)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 0xFF
        out[y] = cur
        prev = cur
    if depth == 16:
        out = out[:, 0::2] * 256 + out[:, 1::2]
        dt = np.uint16
    else:
        dt = np.uint8
    return out.reshape((h, w, ch) if ch == 3 else (h, w)).astype(dt)


@pytest.mark.parametrize("word", [64, 32])
def test_image_modes_match_oracle(ffr_mod, oracle_mod, word):
    text = TF.color_flame(r=3, size=96)
    r = render(ffr_mod, text, 1 << 21, word=word)
    buf = r.read_buffer()
    ffr = ffr_mod
    w, h = 96, 96
    for mode in (ffr.IMG_MONO, ffr.IMG_GRAY, ffr.IMG_COLOR):
        for bits in (8, 16):
            for gamma in (1.0, 2.2):
                pix, info = r.render_image(mode, gamma, bits)
                st, ref, hist, scal = oracle_mod.tonemap(buf, w, h, 4, mode, gamma, bits)
                assert st == 0
                compare(pix, ref, exact=(mode == ffr.IMG_MONO))
                assert (info.hist_min, info.hist_max) == hist
                rel = 1e-15 if word == 64 else 1e-6
                assert abs(info.scale_max - scal[1]) <= rel * scal[1]
                assert abs(info.scale_min - scal[0]) <= rel * max(scal[0], 1.0)
                assert info.width == w and info.height == h
    r.close()


def test_image_of_rectangular_buffer(ffr_mod, oracle_mod):
    """rows in buffer order, x fastest, on a 160 x 90 buffer"""
    text = TF.color_flame(r=3, size=96).replace("[96, 96]", "[160, 90]")
    r = render(ffr_mod, text, 1 << 20)
    buf = r.read_buffer()
    pix, info = r.render_image(ffr_mod.IMG_COLOR, 1.7, 8)
    assert pix.shape == (90, 160, 3) and (info.width, info.height) == (160, 90)
    st, ref, _, _ = oracle_mod.tonemap(buf, 160, 90, 4, 2, 1.7, 8)
    compare(pix, ref)
    r.close()


def test_full_size_headline_image(ffr_mod, oracle_mod):
    """3840 x 2160 random colour flame: the whole image against the oracle"""
    text = read_flame("random_color_3840x2160.json")
    r = render(ffr_mod, text, 1 << 26, batch=8192)
    buf = r.read_buffer()
    for mode, bits in ((ffr_mod.IMG_GRAY, 16), (ffr_mod.IMG_COLOR, 8)):
        pix, _ = r.render_image(mode, 2.2, bits)
        st, ref, _, _ = oracle_mod.tonemap(buf, 3840, 2160, 4, mode, 2.2, bits)
        assert st == 0
        compare(pix, ref)
    r.close()


def test_image_errors(ffr_mod):
    ffr = ffr_mod
    r = ffr.BufferRenderer(TF.color_flame(r=3, size=32), seed=SEED)
    with pytest.raises(ffr.FfrError, match=r"histogram is \(probably\) empty"):
        r.render_image(ffr.IMG_GRAY)
    assert r.render(1 << 14)
    with pytest.raises(ffr.FfrError, match="gamma too small"):
        r.render_image(ffr.IMG_GRAY, gamma=0.0)
    with pytest.raises(ffr.FfrError, match="bits per channel must be 8 or 16"):
        r.render_image(ffr.IMG_GRAY, bits=4)
    with pytest.raises(ffr.FfrError, match="exactly 1 coloring flag"):
        r.render_image(7)
    r.close()
    r = ffr.BufferRenderer(TF.color_flame(r=1, size=32), seed=SEED)
    assert r.render(1 << 14)
    with pytest.raises(ffr.FfrError, match="buffer must use 3 color dimensions"):
        r.render_image(ffr.IMG_COLOR)
    pix, _ = r.render_image(ffr.IMG_GRAY)
    assert pix.shape == (32, 32)
    r.close()
    r = ffr.BufferRenderer(TF.tetrahedron(), seed=SEED)
    with pytest.raises(ffr.FfrError, match="only 2D flames supported"):
        r.render_image(ffr.IMG_GRAY)
    r.close()


def _run(exe, args):
    p = subprocess.run([os.path.join(PKG, exe)] + args, capture_output=True, timeout=120)
    return p.returncode, p.stdout, p.stderr.decode(errors="replace")


@pytest.mark.parametrize("flag,bits", [("-m", 8), ("-g", 8), ("-g", 16), ("-c", 8), ("-c", 16)])
def test_ffr_img_program(ffr_mod, oracle_mod, tmp_path, flag, bits):
    """ffr-buf twice, then ffr-img over both buffers: the PNG holds the
    oracle's image of their sum"""
    f = tmp_path / "flame.json"
    f.write_text(TF.color_flame(r=3, size=80).replace("[80, 80]", "[80, 48]"))
    bufs = []
    for seed in (1, 2):
        b = tmp_path / f"{seed}.buf"
        rc, _, err = _run("ffr-buf", ["-f", str(f), "-o", str(b), "-s", str(1 << 19),
                                      "--seed", str(seed)])
        assert rc == 0, err
        bufs.append(b)
    png = tmp_path / "out.png"
    args = ["-f", str(f), "-i", str(bufs[0]), "-i", str(bufs[1]), "-o", str(png), flag,
            "-b", str(bits), "-y", "2.2"]
    rc, _, err = _run("ffr-img", args)
    assert rc == 0, err
    for line in ("ffr-img version", "--gamma 2.2\n", f"--bits {bits}\n", "buffer: ",
                 "color: 3 dimensions", "adding input file", "processing buffer",
                 "histogram bounds: ", "scaler bounds: "):
        assert line in err, line
    total = np.fromfile(bufs[0], dtype=np.uint64).copy()
    b2 = np.fromfile(bufs[1], dtype=np.uint64)
    cnt = total.reshape(-1, 4)
    cnt[:, 0] += b2.reshape(-1, 4)[:, 0]
    cnt[:, 1:] = (cnt[:, 1:].view(np.float64) + b2.reshape(-1, 4)[:, 1:].view(np.float64)
                  ).view(np.uint64)
    mode = {"-m": 0, "-g": 1, "-c": 2}[flag]
    st, ref, hist, _ = oracle_mod.tonemap(total, 80, 48, 4, mode, 2.2, bits)
    assert st == 0
    assert f"histogram bounds: {hist[0]} {hist[1]}\n" in err
    data = png.read_bytes()
    img = decode_png(data)
    compare(img, ref, exact=(mode == 0))
    try:
        from PIL import Image
    except ImportError:
        return
    if mode == 2 and bits == 16:
        return  # Pillow reads 16-bit RGB as 8-bit
    with Image.open(png) as im:
        arr = np.array(im)
    assert np.array_equal(arr.astype(np.int64), img.astype(np.int64))


def test_ffr_img_program_errors(ffr_mod, tmp_path):
    f = tmp_path / "flame.json"
    f.write_text(TF.color_flame(r=3, size=32))
    b = tmp_path / "a.buf"
    rc, _, err = _run("ffr-buf", ["-f", str(f), "-o", str(b), "-s", "0"])
    assert rc == 0, err
    rc, _, err = _run("ffr-img", ["-f", str(f), "-o", "-", "-g", "-c", "-i", str(b)])
    assert rc == 1 and "must choose exactly 1 coloring flag" in err
    assert "buffer: 32768 bytes" in err
    rc, _, err = _run("ffr-img", ["-f", str(f), "-o", "-", "-g"])
    assert rc == 1 and "ERROR: no input buffer files specified" in err
    rc, out, err = _run("ffr-img", ["-f", str(f), "-o", "-", "-g", "-i", str(b)])
    assert rc == 1 and "histogram is (probably) empty" in err and out == b""
    assert "histogram bounds: 0 0\nscaler bounds: 0 0\n" in err
    short = tmp_path / "short.buf"
    short.write_bytes(b"\0" * 100)
    rc, _, err = _run("ffr-img", ["-f", str(f), "-o", "-", "-g", "-i", str(short)])
    assert rc == 1 and "error reading file" in err


def test_image_into_device_memory(ffr_mod):
    """ffr_render_image_device leaves the same pixels in a caller's device
    buffer, and times the pixel kernel"""
    r = render(ffr_mod, TF.color_flame(r=3, size=64), 1 << 19)
    for mode, bits in ((ffr_mod.IMG_COLOR, 16), (ffr_mod.IMG_MONO, 8)):
        host, _ = r.render_image(mode, 1.5, bits)
        n = r.image_bytes(mode, bits)
        assert n == host.nbytes
        dev = ffr_mod.DeviceMemory(n + 64, fill=0)
        info = r.render_image_device(dev.ptr, n + 64, mode, 1.5, bits)
        assert info.kernel_ms > 0
        raw = dev.to_host()
        got = raw[:n].view(host.dtype).reshape(host.shape)
        assert np.array_equal(got, host)
        assert int(raw[n:].sum()) == 0   # nothing written past the image
        with pytest.raises(ffr_mod.FfrError, match="expected"):
            r.render_image_device(dev.ptr, 16, ffr_mod.IMG_GRAY, 1.0, 8)
        dev.close()
    r.close()


@pytest.mark.parametrize("word", [64, 32])
def test_image_edge_values(ffr_mod, oracle_mod, word):
    """crafted buffers: colour means outside [0, 1] (low bits of a signed
    32-bit conversion, 0 beyond its range), gamma inf and NaN, colour sums in
    empty cells; device and oracle agree (bit for bit in double)"""
    it, ft = (np.uint64, np.float64) if word == 64 else (np.uint32, np.float32)
    w = h = 24
    rng = np.random.default_rng(9)
    counts = rng.integers(0, 50, size=w * h).astype(it)
    counts[::7] = 0
    means = rng.uniform(-2.0, 3.0, size=(w * h, 3))
    means[5] = [1e10, -1e10, 300.0]
    sums = (means * counts[:, None]).astype(ft)
    sums[counts == 0] = 0.25   # colour left in empty cells
    buf = np.zeros((w * h, 4), dtype=it)
    buf[:, 0] = counts
    buf[:, 1:] = sums.view(it)
    buf = buf.reshape(-1)
    r = ffr_mod.BufferRenderer(TF.color_flame(r=3, size=w),
                               precision="fp64" if word == 64 else "fp32", word=word)
    r.write_buffer(buf)
    for gamma in (1.0, 0.3, float("inf"), float("nan")):
        for mode in (0, 1, 2):
            for bits in (8, 16):
                pix, _ = r.render_image(mode, gamma, bits)
                st, ref, _, _ = oracle_mod.tonemap(buf, w, h, 4, mode, gamma, bits)
                assert st == 0
                # float log / pow of the device and of glibc may differ in
                # the last bit; double ones move no pixel of these buffers
                compare(pix, ref, exact=(word == 64 or mode == 0), label=(gamma, mode, bits))
    r.close()


def test_ffr_img_pipe_and_float_build(ffr_mod, oracle_mod, tmp_path):
    """ffr-buf -o - | ffr-img -i -: the buffer streams through both programs;
    and the float build's layout (--word 32) end to end"""
    f = tmp_path / "flame.json"
    f.write_text(TF.color_flame(r=3, size=72).replace("[72, 72]", "[72, 40]"))
    for word in ("64", "32"):
        png = tmp_path / f"pipe{word}.png"
        extra = ["--word", word] + (["--precision", "fp32"] if word == "32" else [])
        buf = subprocess.run([os.path.join(PKG, "ffr-buf"), "-f", str(f), "-o", "-", "-s",
                              str(1 << 18), "--seed", "4"] + extra,
                             capture_output=True, timeout=120)
        assert buf.returncode == 0, buf.stderr.decode()
        img = subprocess.run([os.path.join(PKG, "ffr-img"), "-f", str(f), "-i", "-", "-o",
                              str(png), "-c", "-b", "16", "--word", word],
                             input=buf.stdout, capture_output=True, timeout=120)
        assert img.returncode == 0, img.stderr.decode()
        words = np.frombuffer(buf.stdout, dtype=np.uint64 if word == "64" else np.uint32)
        st, ref, _, _ = oracle_mod.tonemap(words, 72, 40, 4, 2, 1.0, 16)
        assert st == 0
        compare(decode_png(png.read_bytes()), ref, exact=(word == "64"))

"""The benchmark workloads' own shapes against the CPU oracle.

Each GPU path that the benchmark measures is checked here against the
oracle (the reference loop, src/renderers/buffer_renderer.hpp:150-250 and
src/renderers/render_iterator.hpp:52-139, restated on the CPU) on the same
walker streams, not only against another GPU path:

  - the colour flame of the headline workload at 384x216: counts and colour
    sums of the FP64 render against the FP64 oracle, and of the FP32
    render;
  - batched rendering: members of the 256-flame batch set at their real
    resolutions (256^2, 512^2, 1024^2, 2048^2) with the sample count cut to
    what the oracle renders in a second, rendered together in one batch;
    the same flames with every variation made linear, bit for bit;
  - serving: jobs of four chunks mixed with other jobs on the render
    server, chunk by chunk against the oracle (a linear 2-D job and a 3-D
    float-build job bit for bit, the headline colour flame within the L1
    bar);
  - the full 256-flame batch once (3.6e11 samples): every flame's counts sum
    to its samples plotted, and one member of each resolution equals its
    single render;
  - the precision policy: flames it keeps in FP32 are within L1 1e-3 of
    the FP64 oracle.

Bars (the north star's): linear flames bit-exact counts and exact
statistics, colour sums within 1e-12 relative (atomic order); flames with
transcendental variations at normalised-histogram L1 <= 1e-3 against the
FP64 oracle, per-channel colour L1 <= 1e-3, and colour sums within 1e-5
relative in cells whose counts agree.  The xform selection draws and the
colour recurrence do not depend on the point, so samples iterated are exact
for every flame, and xform counts whenever both sides draw in the same
precision.
"""

import json
import os
import re
import time

import numpy as np
import pytest

from conftest import FLAMES, read_flame

pytestmark = pytest.mark.gpu

SETTLE = 53
THREADS = min(16, os.cpu_count() or 8)
F64_U64, F32_U32, F32_U64 = 0, 1, 2


def load(text):
    return json.loads(re.sub(r"^\s*//.*$", "", text, flags=re.M))


def batch_flame(name):
    with open(os.path.join(FLAMES, "batch", name)) as f:
        return f.read()


def pick_batch_flames():
    """the first flame of each resolution in the batch set"""
    out = {}
    for name in sorted(os.listdir(os.path.join(FLAMES, "batch"))):
        d = load(batch_flame(name))
        out.setdefault(d["size"][0], name)
    assert sorted(out) == [256, 512, 1024, 2048]
    return [out[k] for k in sorted(out)]


def linearised(text):
    """the same flame with every variation replaced by linear of the same
    weight: contractive affine maps, no transcendental arithmetic"""
    d = load(text)
    for x in d["xforms"]:
        x["variations"] = [{"name": "linear", "weight": v["weight"]} for v in x["variations"]]
    return json.dumps(d)


def l1(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def colour_report(gc, gcol, oc, ocol):
    """(per-channel normalised L1, cells with equal non-zero counts, share of
    those whose colour sums agree within 1e-5 relative, worst relative
    difference among them)"""
    ch = [float(np.abs(gcol[:, k] - ocol[:, k]).sum() / max(np.abs(ocol[:, k]).sum(), 1e-300))
          for k in range(ocol.shape[1])]
    eq = (gc == oc) & (oc > 0)
    g, o = gcol[eq].astype(np.float64), ocol[eq].astype(np.float64)
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-300)
    rel[(g == 0) & (o == 0)] = 0.0
    per_cell = rel.max(axis=1) if rel.size else np.zeros(0)
    ok = float((per_cell <= 1e-5).mean()) if per_cell.size else 1.0
    return ch, int(eq.sum()), ok, float(per_cell.max()) if per_cell.size else 0.0


def oracle_render(oracle, text, mode, samples, batch, seed, settle=SETTLE, walker_base=0,
                  buffer=None):
    f = oracle.OracleFlame(text, mode, settle=settle)
    return f, f.render(samples, batch, seed=seed, walker_base=walker_base, threads=THREADS,
                       buffer=buffer)


def assert_exact(r, ref, dims, colour_tol=1e-12):
    w = r.read_buffer()
    assert np.array_equal(r.counts(w), ref.counts)
    if r.color_dims:
        gc = r.colors(w)
        assert np.allclose(gc, ref.colors, rtol=colour_tol,
                           atol=colour_tol * max(1.0, float(np.abs(ref.colors).max())))
    assert r.samples_iterated == ref.s_iter and r.samples_plotted == ref.s_plot
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    ext = r.point_extremes
    for i in range(dims):
        assert ext[i] == (ref.pt_min[i], ref.pt_max[i])


def assert_within_bars(r, ref, label, colour=True, same_precision=True):
    """transcendental flames: L1 and colour bars, exact iteration counts, and
    exact xform counts when both sides draw in the same precision (FP32
    turns a word into a 24-bit number and compares it with float cumulative
    weights, FP64 a 53-bit one with double weights, so a draw near a weight
    boundary may pick another xform); returns the L1.

    Colour sums of a cell agree when the same samples landed there.  Where a
    trajectory parted from the oracle's, a cell may hold the same count of
    other samples; FP32 trajectories against the FP64 oracle part more often
    than FP64 ones, so their share of equal-count cells within 1e-5 is held
    to 99 % (99.9 % for the same precision)."""
    w = r.read_buffer()
    gc = r.counts(w)
    d = l1(gc, ref.counts)
    assert r.samples_iterated == ref.s_iter
    if same_precision:
        assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    else:
        xd, xr = r.xform_distribution().astype(float), ref.xf_dist.astype(float)
        assert np.abs(xd - xr).sum() <= 1e-4 * xr.sum()
    assert abs(r.samples_plotted - ref.s_plot) <= 1e-3 * ref.s_plot
    msg = f"{label}: L1 {d:.2e}"
    if colour and r.color_dims:
        ch, n_eq, ok, worst = colour_report(gc, r.colors(w), ref.counts, ref.colors)
        msg += (f", colour L1 {max(ch):.2e}, {n_eq} equal-count cells, "
                f"{100 * ok:.3f} % within 1e-5 (worst {worst:.1e})")
        assert max(ch) <= 1e-3, msg
        assert ok >= (0.999 if same_precision else 0.99), msg
    print(msg, flush=True)
    assert d <= 1e-3, msg
    return d


# --------------------------------------------------- headline colour flame

def test_headline_colour_flame_against_oracle(ffr_mod, oracle_mod):
    """the headline colour flame (6 xforms of transcendental variations,
    three colour dimensions) at 384x216, 2^21 samples"""
    text = read_flame("random_color_3840x2160.json").replace("3840,\n  2160", "384,\n  216")
    assert '"size": [\n  384' in text
    samples, batch, seed = 1 << 21, 4096, 0x5EED0002
    _, ref = oracle_render(oracle_mod, text, F64_U64, samples, batch, seed, settle=-1)
    assert ref.success and ref.colors.shape[1] == 3
    r64 = ffr_mod.BufferRenderer(text, precision="fp64", seed=seed, batch_size=batch)
    assert r64.render(samples)
    assert_within_bars(r64, ref, "headline 384x216 FP64 vs oracle FP64")
    r64.close()
    # FP32 trajectories on the same streams (settle 53 as FP64, so both draw
    # the same words), against the same FP64 oracle render
    r32 = ffr_mod.BufferRenderer(text, precision="fp32", seed=seed, batch_size=batch, settle=53)
    assert r32.render(samples)
    assert_within_bars(r32, ref, "headline 384x216 FP32 vs oracle FP64", same_precision=False)
    r32.close()


# --------------------------------------------------------------- batching

@pytest.mark.parametrize("kind", ["transcendental", "linear"])
def test_batch_members_at_real_resolutions(ffr_mod, oracle_mod, kind):
    names = pick_batch_flames()
    texts = [batch_flame(n) for n in names]
    if kind == "linear":
        texts = [linearised(t) for t in texts]
    samples, batch = [1 << 22, 1 << 22, 3 << 21, 1 << 23], 4096
    seeds = [0x5EED0004 + 17 * i for i in range(len(texts))]
    precs = ["fp64"] * 4 + (["fp32"] * 4 if kind == "linear" else [])
    all_texts = texts + (texts if kind == "linear" else [])
    all_samples = samples + (samples if kind == "linear" else [])
    all_seeds = seeds + (seeds if kind == "linear" else [])
    rs = [ffr_mod.BufferRenderer(t, precision=p, seed=s, batch_size=batch, settle=SETTLE)
          for t, p, s in zip(all_texts, precs, all_seeds)]
    with ffr_mod.Batch(rs) as b:
        print(f"batch of {len(rs)} ({kind}): {names}", flush=True)
        assert all(b.render(all_samples))
        assert b.last_stats.graph == 1
    for i, (r, t, p, n, s) in enumerate(zip(rs, all_texts, precs, all_samples, all_seeds)):
        mode = F64_U64 if p == "fp64" else F32_U64
        t0 = time.perf_counter()
        _, ref = oracle_render(oracle_mod, t, mode, n, batch, s)
        assert ref.success
        label = f"{names[i % 4]} {r.cells} cells {p} {kind}"
        if kind == "linear":
            assert_exact(r, ref, 2)
            print(f"{label}: bit-exact ({time.perf_counter() - t0:.1f} s oracle)", flush=True)
        else:
            assert_within_bars(r, ref, label)
        r.close()


# ---------------------------------------------------------------- serving

def test_served_jobs_against_oracle(ffr_mod, oracle_mod):
    """three checked jobs of four chunks on one server, mixed with filler
    jobs, passes short enough that every chunk is sliced (the largest job's
    chunks are 367 walkers of 2048 samples, six passes each)"""
    names = pick_batch_flames()
    lin = linearised(batch_flame(names[2]))                       # 1024^2
    headline = read_flame("random_color_3840x2160.json").replace("3840,\n  2160", "384,\n  216")
    with open(os.path.join(FLAMES, "serve3d", "flame_3000.json")) as f:
        vol = f.read()                                            # 256^3 u32
    checked = [
        # (text, precision, word, mode, samples, seed, exact)
        (lin, "fp64", 64, F64_U64, 3_000_001, 71, True),
        (vol, "fp32", 32, F32_U32, 2_500_003, 72, True),
        (headline, "fp64", 64, F64_U64, 1 << 21, 73, False),
    ]
    fillers = [(batch_flame(n), "fp32", 64, 1_000_000 + 7 * k, 80 + k)
               for k, n in enumerate(names[:2] * 2)]
    chunks, batch = 4, 2048
    # the smallest pass target: each job gets the smallest slice (64
    # walkers) per pass, so every chunk spreads over several passes
    with ffr_mod.Server(pass_ms=0.001) as srv:
        ids = []
        for k, (text, prec, word, n, seed) in enumerate(fillers):
            srv.submit(text, n, chunks, precision=prec, word=word, seed=seed, batch_size=batch)
            if k == 1:
                for text2, prec2, word2, _, n2, seed2, _ in checked:
                    ids.append(srv.submit(text2, n2, chunks, precision=prec2, word=word2,
                                          seed=seed2, batch_size=batch))
        bufs = []
        for j in ids:
            info = srv.wait(j, timeout=120)
            assert info.state == ffr_mod.JOB_DONE, srv.job_error(j)
            bufs.append((srv.read(j, word=64 if j != ids[1] else 32), info))
        st = srv.stats()
        print(f"server: {st.passes} passes, {st.done} jobs", flush=True)
        assert st.passes >= 4 * 6
    for (text, prec, word, mode, n, seed, exact), (got, info) in zip(checked, bufs):
        assert info.batch_size == batch and info.chunks_done == chunks
        f = oracle_mod.OracleFlame(text, mode, settle=-1)
        buf = f.new_buffer()
        per, base = n // chunks, 0
        for c in range(chunks):
            m = n - per * (chunks - 1) if c == chunks - 1 else per
            ref = f.render(m, batch, seed=seed, walker_base=base, threads=THREADS, buffer=buf)
            assert ref.success
            base += (m + batch - 1) // batch
        cw = f.cell_words
        gc = got.reshape(-1, cw)[:, 0]
        oc = buf.reshape(-1, cw)[:, 0]
        label = f"served {f.dims}-D {prec} job, {f.cells} cells"
        if exact:
            assert np.array_equal(got, buf) if cw == 1 else np.array_equal(gc, oc), label
            if cw > 1:
                gcol = got.reshape(-1, cw)[:, 1:].view(np.float64)
                ocol = buf.reshape(-1, cw)[:, 1:].view(np.float64)
                assert np.allclose(gcol, ocol, rtol=1e-12, atol=1e-12 * float(ocol.max()))
            print(f"{label}: bit-exact counts over {chunks} chunks", flush=True)
        else:
            d = l1(gc, oc)
            gcol = got.reshape(-1, cw)[:, 1:].view(np.float64)
            ocol = buf.reshape(-1, cw)[:, 1:].view(np.float64)
            ch, n_eq, ok, worst = colour_report(gc, gcol, oc, ocol)
            msg = (f"{label}: L1 {d:.2e}, colour L1 {max(ch):.2e}, {100 * ok:.3f} % of "
                   f"{n_eq} equal-count cells within 1e-5 (worst {worst:.1e})")
            print(msg, flush=True)
            assert d <= 1e-3 and max(ch) <= 1e-3 and ok >= 0.999, msg


# ------------------------------------------------------- the whole batch

def test_full_batch_set_properties(ffr_mod):
    """the 256-flame batch at 1000 samples per cell, FP32, once: each
    flame's counts sum to its samples plotted and it iterated exactly its
    samples; one member of each resolution equals its own single render"""
    names = sorted(n for n in os.listdir(os.path.join(FLAMES, "batch")) if n.endswith(".json"))
    texts = [batch_flame(n) for n in names]
    cells = [load(t)["size"][0] * load(t)["size"][1] for t in texts]
    samples = [1000 * c for c in cells]
    seed0 = 0x5EED0004
    rs = [ffr_mod.BufferRenderer(t, precision="fp32", seed=seed0 + i, settle=SETTLE)
          for i, t in enumerate(texts)]
    with ffr_mod.Batch(rs) as b:
        t0 = time.perf_counter()
        assert all(b.render(samples))
        dt = time.perf_counter() - t0
        print(f"256 flames, {sum(samples):.3e} samples in {dt:.1f} s "
              f"({sum(samples) / dt / 1e9:.1f} G samples/s)", flush=True)
    for r, n in zip(rs, samples):
        s, _, _ = r.histogram_stats()
        assert r.samples_iterated == n
        assert s == r.samples_plotted
        assert r.samples_plotted <= n
    picks = {}
    for i, c in enumerate(cells):
        picks.setdefault(c, i)
    for c, i in sorted(picks.items()):
        one = ffr_mod.BufferRenderer(texts[i], precision="fp32", seed=seed0 + i, settle=SETTLE)
        assert one.render(samples[i])
        a, w = rs[i].read_buffer(), one.read_buffer()
        assert np.array_equal(one.counts(a), one.counts(w)), names[i]
        ca, cw_ = one.colors(a), one.colors(w)
        assert np.allclose(ca, cw_, rtol=1e-12, atol=1e-12 * float(np.abs(cw_).max()))
        assert np.array_equal(rs[i].xform_distribution(), one.xform_distribution())
        assert rs[i].point_extremes == one.point_extremes
        print(f"{names[i]} ({c} cells): batch member equals its single render", flush=True)
        one.close()
    for r in rs:
        r.close()


# ------------------------------------------------------ precision policy

def test_policy_fp32_flames_against_fp64_oracle(ffr_mod, oracle_mod):
    """32 flames of the batch set at 64x64 (1000 samples per cell): the
    policy's choices, then each flame in its chosen precision on the GPU
    against the FP64 oracle on the same streams"""
    names = sorted(n for n in os.listdir(os.path.join(FLAMES, "batch")) if n.endswith(".json"))
    names = names[::8]
    texts = []
    for n in names:
        d = load(batch_flame(n))
        d["size"] = [64, 64]
        texts.append(json.dumps(d))
    n = 1000 * 64 * 64
    seeds = [0x5EED0004 + i for i in range(len(texts))]
    precs, probe = ffr_mod.choose_precisions(texts, [n] * len(texts), seeds)
    worst = 0.0
    for t, p, s, nm, pr in zip(texts, precs, seeds, names, probe):
        r = ffr_mod.BufferRenderer(t, precision=p, seed=s, settle=SETTLE)
        assert r.render(n)
        _, ref = oracle_render(oracle_mod, t, F64_U64, n, r.last_stats.batch_size, s)
        d = l1(r.counts(), ref.counts)
        worst = max(worst, d)
        print(f"{nm}: policy {p} (probe L1 {pr:.1e}), L1 vs FP64 oracle {d:.2e}", flush=True)
        assert d <= 1e-3, (nm, p, d)
        r.close()
    assert precs.count("fp32") >= len(texts) // 2
    print(f"policy: {precs.count('fp32')} FP32 / {precs.count('fp64')} FP64, worst L1 "
          f"{worst:.2e}", flush=True)

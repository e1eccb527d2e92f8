"""The render server on the GPU: jobs admitted within an HBM budget, buffers
kept on the device across chunks, continuous batching of every running job
into shared passes.  Each job's result must equal the same job rendered
alone, chunk by chunk, with BufferRenderer (counts and statistics exact,
colours within 1e-12 relative), whatever else shared its passes.
"""

import json
import threading
import time

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu


def job_mix():
    """(flame, precision, word, samples): several kinds and sizes."""
    c2 = read_flame("random_color_3840x2160.json").replace("3840,\n  2160", "160,\n  90")
    assert '"size": [\n  160' in c2
    sier = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[96, 96]")
    return [
        (c2, "fp32", 64, 1_500_000),
        (sier, "fp64", 64, 700_003),
        (TF.color_flame(3), "fp64", 64, 900_000),
        (TF.tetrahedron(), "fp32", 32, 400_000),
        (c2, "fp64", 64, 600_000),
        (TF.color_flame(4, size=80, seed=7), "fp32", 64, 350_001),
        (TF.one_d(), "fp32", 64, 120_000),
        (sier, "fp32", 32, 1_000_000),
    ]


def alone(ffr, text, precision, word, samples, chunks, seed, batch):
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=seed, batch_size=batch)
    per = samples // chunks
    for c in range(chunks):
        n = samples - per * (chunks - 1) if c == chunks - 1 else per
        assert r.render(n)
    return r


def compare(ffr, srv, job, text, precision, word, samples, chunks, seed):
    info = srv.job(job)
    assert info.state == ffr.JOB_DONE, srv.job_error(job)
    assert info.chunks_done == chunks and info.samples_done == samples
    assert 0 <= info.submit_s <= info.admit_s <= info.first_chunk_s <= info.done_s
    got = srv.read(job, word)
    r = alone(ffr, text, precision, word, samples, chunks, seed, info.batch_size)
    want = r.read_buffer()
    assert np.array_equal(r.counts(got), r.counts(want))
    if r.color_dims:
        tol = 1e-12 if word == 64 else 2e-6
        gc, wc = r.colors(got), r.colors(want)
        assert np.allclose(gc, wc, rtol=tol, atol=tol * max(1.0, float(np.abs(wc).max())))
    r.close()


@pytest.mark.parametrize("pass_ms", [0.001, 40.0])
def test_jobs_equal_alone_in_any_mix(ffr_mod, pass_ms):
    """A tiny pass target gives each job the smallest slice (64 walkers), so
    every chunk spreads over many passes; staggered arrivals change the mix
    of each pass."""
    mix = job_mix()
    chunks = 4
    with ffr_mod.Server(pass_ms=pass_ms) as srv:
        jobs = []
        for i, (text, prec, word, n) in enumerate(mix):
            jobs.append(srv.submit(text, n, chunks, precision=prec, word=word, seed=50 + i))
            time.sleep(0.003 * (i % 3))
        for j in jobs:
            srv.wait(j, timeout=120)
        st = srv.stats()
        print(f"pass_ms {pass_ms}: {st.passes} passes, {st.samples} samples", flush=True)
        assert st.done == len(mix) and st.failed == 0
        assert st.samples == sum(n for _, _, _, n in mix)
        if pass_ms < 1:
            # the longest job: 1.5e6 samples, 4 chunks of 366 walkers of 1024,
            # 64 walkers a pass
            assert st.passes >= 4 * 5
        for i, ((text, prec, word, n), j) in enumerate(zip(mix, jobs)):
            compare(ffr_mod, srv, j, text, prec, word, n, chunks, 50 + i)


def test_admission_within_budget(ffr_mod):
    text = TF.color_flame(3)  # 128 x 128 x 4 words x 8 bytes
    bytes_per_job = 128 * 128 * 4 * 8
    budget = int(2.5 * bytes_per_job)
    with ffr_mod.Server(hbm_budget=budget, pass_ms=2.0) as srv:
        jobs = [srv.submit(text, 400_000, 2, precision="fp64", seed=i) for i in range(6)]
        done = []

        def drain():
            for j in jobs:
                srv.wait(j, timeout=120)
                done.append(srv.job(j))
                srv.release(j)

        t = threading.Thread(target=drain)
        t.start()
        t.join(timeout=180)
        assert not t.is_alive()
        st = srv.stats()
        assert st.done == 6 and st.peak_bytes <= budget and st.bytes_held == 0
        # at most two jobs held at once: job k waits for job k-2's release
        for k in range(2, 6):
            assert done[k].admit_s >= done[k - 2].done_s
    with pytest.raises(ffr_mod.FfrError):
        with ffr_mod.Server(hbm_budget=bytes_per_job // 2) as srv:
            srv.submit(text, 1000, 1)


def test_starting_buffer_is_added(ffr_mod):
    """ffr-buf -i: a job may start from an existing buffer."""
    text = TF.color_flame(3)
    base = ffr_mod.BufferRenderer(text, seed=9, batch_size=1000)
    assert base.render(300_000)
    init = base.read_buffer()
    with ffr_mod.Server(pass_ms=2.0) as srv:
        j = srv.submit(text, 200_000, 2, precision="fp64", seed=11, batch_size=1000, init=init)
        srv.wait(j, timeout=60)
        got = srv.read(j)
    r = alone(ffr_mod, text, "fp64", 64, 200_000, 2, 11, 1000)
    assert np.array_equal(r.counts(got), r.counts(r.read_buffer()) + base.counts(init))
    r.close()
    base.close()


def test_partial_reads_grow(ffr_mod):
    text = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[128, 128]")
    with ffr_mod.Server(pass_ms=0.001) as srv:
        j = srv.submit(text, 4_000_000, 4, precision="fp64", seed=3, batch_size=1000)
        seen = []
        while srv.job(j).state != ffr_mod.JOB_DONE:
            info = srv.job(j)
            if info.state == ffr_mod.JOB_RUNNING:
                seen.append((info.samples_done, int(srv.read(j)[::1].sum())))
            time.sleep(0.001)
        final = int(srv.read(j).sum())
    assert final == 4_000_000
    # a read between passes holds exactly the samples done at that point
    # (or more: a pass may finish between the two calls)
    for done_before, total in seen:
        assert done_before <= total <= final


def test_rejections(ffr_mod):
    with ffr_mod.Server() as srv:
        with pytest.raises(ffr_mod.FfrError, match="5D not supported|dimensions"):
            srv.submit(json.dumps({"dimensions": 5}), 1000, 1)
        with pytest.raises(ffr_mod.FfrError):
            srv.submit(TF.fern(), 1000, 0)
        j = srv.submit(TF.expanding(), 100_000, 1, precision="fp64", bad_value_limit=3)
        info = srv.wait(j, timeout=60)
        assert info.state == ffr_mod.JOB_FAILED and info.status == 4
        assert "contractive" in srv.job_error(j)

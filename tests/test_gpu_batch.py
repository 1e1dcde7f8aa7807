"""Batched multi-job wavefront against single renders of the same jobs.

Bar: after every step each job's queue lengths and started samples equal those of a single
render with the same scene, camera, slot count and seed; the job's slots hold bit-identical
state (RNG, rays, paths, hits, shading, medium lists) and its queue runs list the same slots
in the same order; per-pixel sample counts are equal and radiance sums agree within 1e-4
relative (float atomics add in arbitrary order). Sample budgets start every pixel exactly
spp times and then retire the job's slots.
"""
import numpy as np
import pytest

from raccoongpurender_amd import _lib
from raccoongpurender_amd import render_scenes as R
from tests import wf_helpers as W

pytestmark = pytest.mark.gpu


def _device_scene(ctx, sd):
    bvh = _lib.BVH(sd["points"], sd["indices"])
    return dict(bvh=bvh, scene=_lib.Scene(ctx, bvh), mats=_lib.Materials(ctx, sd["materials"]),
                env=_lib.Envmap(ctx, sd["env"]), camera=W.camera(sd["camera"]))


def _jobs(ctx, specs):
    out = []
    for sd, slots, spp, seed in specs:
        d = _device_scene(ctx, sd)
        j = _lib.BatchJob(d["scene"], d["mats"], d["env"], d["camera"], slots, spp=spp, seed=seed)
        j.keep = d
        out.append(j)
    return out


def _single(ctx, job, **kw):
    wf = _lib.Wavefront(ctx, job.scene, job.materials, job.envmap, job.slots, seed=job.seed, **kw)
    wf.initialize(job.camera)
    return wf


def _job_items(batch_state, totals, first, slots):
    items = batch_state["items"]
    # the job's run inside each global queue: the live entries that fall in its slot range
    out = []
    for q in range(7):
        row = items[q, :totals[q]]
        sel = row[(row >= first) & (row < first + slots)]
        out.append(sel - first)
    return out


SPECS = [
    (R.cornell_box(32, 32, segments=16, rings=8), 2048, 0, 11),
    (R.material_test_scene(32, 24, volume=True), 1536, 0, 12),
    (R.furnace_scene("volume", 16, 16, value=0.5), 512, 0, 13),
]


def test_batch_jobs_match_single_renders(ctx):
    jobs = _jobs(ctx, SPECS)
    batch = _lib.Batch(ctx, jobs)
    singles = [_single(ctx, j) for j in jobs]
    for step in range(12):
        batch.step()
        for s in singles:
            s.step()
        bs = batch.state()
        totals, _ = batch.queue_counts()
        for k, (j, s) in enumerate(zip(jobs, singles)):
            cj, started = batch.job_queue_counts(k)
            cs, npx = s.queue_counts()
            assert np.array_equal(cj, cs), (step, k, cj, cs)
            assert started == npx, (step, k, started, npx)
            ss = s.state()
            sl = batch.job_slots(k)
            sub = {key: bs[key][sl] for key in ("rng", "rays", "paths", "ext", "shade", "vol")}
            mm = W.state_mismatches(sub, ss)
            assert all(v == 0 for v in mm.values()), (step, k, mm)
            runs = _job_items(bs, totals, sl.start, j.slots)
            for q in range(7):
                assert np.array_equal(runs[q], ss["items"][q, :cs[q]]), (step, k, q)
        assert np.array_equal(totals, sum(batch.job_queue_counts(k)[0] for k in range(len(jobs))))
    for k, (j, s) in enumerate(zip(jobs, singles)):
        rb, nb = batch.job_accumulation(k)
        rs, ns = s.accumulation()
        assert np.array_equal(nb, ns), k
        rel = np.abs(rb - rs) / np.maximum(np.abs(rs), 1e-30)
        assert np.all((rel <= 1e-4) | (np.abs(rb - rs) <= 1e-6)), (k, rel.max())
        assert batch.job_stat(k) == int(ns.sum())


def test_sample_budgets_finish_every_pixel_exactly(ctx):
    specs = [(R.cornell_box(24, 16, segments=16, rings=8), 512, 3, 21),
             (R.material_test_scene(16, 16, volume=True), 768, 5, 22),
             (R.furnace_scene("lambertian", 8, 8, value=0.5), 256, 16, 23)]
    jobs = _jobs(ctx, specs)
    batch = _lib.Batch(ctx, jobs, clamp=1e30)
    steps = batch.run(max_steps=4000, check_every=8)
    print(f"batch finished in {steps} steps", flush=True)
    totals, _ = batch.queue_counts()
    assert int(totals.sum()) == 0 and steps < 4000
    for k, j in enumerate(jobs):
        _, cnt = batch.job_accumulation(k)
        c, started = batch.job_queue_counts(k)
        assert started == j.npixels * j.spp and int(c.sum()) == 0
        assert np.all(cnt == j.spp), (k, np.unique(cnt))
        img = batch.job_tonemap(k)
        assert img.shape == (j.npixels, 4) and np.all(img[:, 3] == 255)


def test_batch_graph_replays_and_launch_shapes_agree(ctx):
    jobs = _jobs(ctx, SPECS[:2])
    eager = _lib.Batch(ctx, jobs)
    graphed = _lib.Batch(ctx, jobs)
    shaped = _lib.Batch(ctx, jobs)
    shaped.set_launch("stackless", 64, 96)  # same traversal order, other workgroup shapes
    graphed.graph_capture(4)
    for _ in range(3):
        eager.step(4)
        graphed.graph_launch(1)
        shaped.step(4)
        for k in range(len(jobs)):
            ce = eager.job_queue_counts(k)
            assert np.array_equal(ce[0], graphed.job_queue_counts(k)[0]) and ce[1] == graphed.job_queue_counts(k)[1]
            assert np.array_equal(ce[0], shaped.job_queue_counts(k)[0])
    for k in range(len(jobs)):
        re, ne = eager.job_accumulation(k)
        for other in (graphed, shaped):
            ro, no = other.job_accumulation(k)
            assert np.array_equal(ne, no)
            assert np.allclose(ro, re, rtol=1e-4, atol=1e-6)
    ms = eager.autotune(None, warmup=1, steps=2)
    assert ms > 0.0


def _furnace_batch(ctx, guided, seed=5):
    from raccoongpurender_amd import guide as G
    specs = [(R.furnace_scene("lambertian", 32, 32, value=0.5, fill=True), 8192, 48, seed),
             (R.furnace_scene("lambertian", 24, 16, value=0.5, fill=True, flip=True), 4096, 64, seed + 1)]
    batch = _lib.Batch(ctx, _jobs(ctx, specs), clamp=1e30)
    gd = None
    if guided:
        gd = G.Guide(ctx, capacity=1 << 14, minibatch=4096, seed=seed)
        G.attach(batch, gd, sampling=True, training=True)
    return batch, gd


def test_guided_batch_samples_carry_the_job_scene(ctx):
    # one network for the batch: each training sample names its job's scene (class 1 or 2 in
    # order of first use) and its position is normalised to that scene's box
    batch, gd = _furnace_batch(ctx, guided=True)
    batch.step(3)
    s = gd.buffered_samples()
    cls, cnt = np.unique(s["scene_class"], return_counts=True)
    print(f"{len(s)} samples, classes {dict(zip(cls.tolist(), cnt.tolist()))}", flush=True)
    assert set(cls.tolist()) == {1.0, 2.0}
    assert (s["p"] >= 0).all() and (s["p"] <= 1).all()
    assert np.isfinite(s["weight"]).all() and (s["weight"] > 0).all()


@pytest.mark.parametrize("guided", [False, True])
def test_guided_batch_furnace_is_exact(ctx, guided):
    # Two lambertian furnaces (R = 0.5 under a sky of 0.5: every sample's expectation is 0.25)
    # in one guided batch that trains while it renders: each job's mean must stay at 0.25.
    batch, gd = _furnace_batch(ctx, guided)
    steps = batch.run(max_steps=4000, check_every=4)
    for k, job in enumerate(batch.jobs):
        rgb, cnt = batch.job_accumulation(k)
        assert (cnt == job.spp).all()
        m = rgb[:, 0] / cnt
        se = m.std(ddof=1) / np.sqrt(m.size)
        z = abs(m.mean() - 0.25) / se
        print(f"guided={guided} job {k}: {steps} steps, mean {m.mean():.6f} se {se:.3g} z={z:.2f}", flush=True)
        assert z < 5.0

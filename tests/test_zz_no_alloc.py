"""rgpu_wavefront_job_tonemap_noalloc against rgpu_wavefront_job_tonemap.

Bar: bit-identical RGBA8 images through the C ABI, on two calls in a row where the second
job is the smaller one (the static staging keeps the first call's contents), and a job above
1024 x 1024 pixels refused with RGPU_ERR_INVALID_ARG.
"""
import numpy as np
import pytest

from raccoongpurender_amd import _lib
from raccoongpurender_amd import render_scenes as R
from tests import wf_helpers as W

pytestmark = pytest.mark.gpu


def _job(ctx, sd, slots, seed, camera=None):
    bvh = _lib.BVH(sd["points"], sd["indices"])
    scene = _lib.Scene(ctx, bvh)
    mats = _lib.Materials(ctx, sd["materials"])
    env = _lib.Envmap(ctx, sd["env"])
    j = _lib.BatchJob(scene, mats, env, camera or W.camera(sd["camera"]), slots, seed=seed)
    j.keep = (bvh, scene, mats, env)
    return j


def test_noalloc_tonemap_equals_allocating_tonemap(ctx):
    big = _job(ctx, R.cornell_box(48, 40, segments=16, rings=8), 2048, 21)
    small = _job(ctx, R.material_test_scene(16, 12, volume=True), 512, 22)
    batch = _lib.Batch(ctx, [big, small])
    batch.step(12)
    for job in (0, 1):  # the larger job first, then the smaller one
        want = batch.job_tonemap(job)
        got = batch.job_tonemap_noalloc(job)
        print(f"job {job}: {batch.jobs[job].npixels} pixels", flush=True)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
        assert np.any(want[:, :3] > 0), "the image should not be black after 12 steps"
    # a second round after more steps: the staging holds the smaller image's bytes now
    batch.step(4)
    for job in (0, 1):
        assert np.array_equal(batch.job_tonemap_noalloc(job), batch.job_tonemap(job))
    batch.close()


def test_noalloc_tonemap_refuses_jobs_above_the_staging(ctx):
    sd = R.material_test_scene(16, 12, volume=False)
    cam = W.camera(dict(sd["camera"], width=1025, height=1024))
    job = _job(ctx, sd, 256, 5, camera=cam)
    batch = _lib.Batch(ctx, [job])
    with pytest.raises(_lib.RgpuError) as ei:
        batch.job_tonemap_noalloc(0)
    assert "1024x1024" in str(ei.value)
    assert batch.job_tonemap(0).shape == (1025 * 1024, 4)
    batch.close()

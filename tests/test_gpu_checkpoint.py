"""Checkpoint / resume of a single render (rgpu_wavefront_checkpoint / rgpu_wavefront_resume).

The reference keeps a lane's state only in device buffers and re-initialises on restart
(common/wavefront_path_tracing.hpp:592-604). The snapshot holds everything a lane carries from
one step to the next. Bar: a lane resumed from it (a new lane or the same one rewound) gives,
after the same number of further steps, the uninterrupted lane's slot states bit for bit, the
same queue counts and pixel cursor, the same per-pixel sample counts, and radiance sums within
1e-4 relative (the accumulation adds with float atomics in arbitrary order).
"""
import numpy as np
import pytest

from raccoongpurender_amd import _lib, render_scenes as R
from tests import wf_helpers as W

pytestmark = pytest.mark.gpu


def _assert_same(g, h):
    ca, pa = g.queue_counts()
    cb, pb = h.queue_counts()
    assert np.array_equal(ca, cb) and pa == pb, (ca, cb, pa, pb)
    mm = W.state_mismatches(g.state(), h.state())
    assert all(v == 0 for v in mm.values()), mm
    ra, na = g.accumulation()
    rb, nb = h.accumulation()
    assert np.array_equal(na, nb)
    rel = np.abs(ra - rb) / np.maximum(np.abs(rb), 1e-30)
    assert np.all((rel <= 1e-4) | (np.abs(ra - rb) <= 1e-6)), rel.max()


def test_resume_into_a_new_lane_continues_exactly(ctx):
    print("case checkpoint: material scene, new lane", flush=True)
    sd = R.material_test_scene(32, 24)
    n = 4096
    a = W.gpu_lane(ctx, sd, n)
    a.step(5)
    blob = a.checkpoint()
    a.step(6)
    b = W.gpu_lane(ctx, sd, n)
    b.step(2)  # a state of its own that the resume must replace
    b.resume(blob)
    b.step(6)
    _assert_same(a, b)


def test_resume_rewinds_and_restores_the_resolution(ctx):
    print("case checkpoint: Cornell box, rewind and other resolution", flush=True)
    sd = R.cornell_box(24, 24, segments=16, rings=8)
    n = 2048
    a = W.gpu_lane(ctx, sd, n)
    a.step(3)
    blob = a.checkpoint()
    a.step(4)
    ref_state, (ref_rgb, ref_cnt) = a.state(), a.accumulation()
    ref_counts = a.queue_counts()
    # the same lane, rewound
    a.resume(blob)
    a.step(4)
    assert W.state_mismatches(a.state(), ref_state) == {k: 0 for k in W.state_mismatches(ref_state, ref_state)}
    assert np.array_equal(a.accumulation()[1], ref_cnt)
    # a lane initialised at another resolution takes the snapshot's
    cam = dict(sd["camera"], width=16, height=8)
    c = W.gpu_lane(ctx, dict(sd, camera=cam), n)
    c.resume(blob)
    assert c.npixels == 24 * 24
    c.step(4)
    assert np.array_equal(c.queue_counts()[0], ref_counts[0])
    rgb, cnt = c.accumulation()
    assert np.array_equal(cnt, ref_cnt)
    assert np.allclose(rgb, ref_rgb, rtol=1e-4, atol=1e-6)


def test_checkpoint_refusals(ctx):
    print("case checkpoint: refused snapshots", flush=True)
    sd = R.cornell_box(16, 16, segments=8, rings=4)
    a = W.gpu_lane(ctx, sd, 1024)
    a.step(2)
    blob = a.checkpoint()
    with pytest.raises(_lib.RgpuError):
        a.resume(blob[:-1])  # truncated
    bad = blob.copy()
    bad[0] ^= 0xFF  # not a checkpoint
    with pytest.raises(_lib.RgpuError):
        a.resume(bad)
    other = W.gpu_lane(ctx, sd, 2048)  # another path count
    with pytest.raises(_lib.RgpuError):
        other.resume(blob)
    wide = W.gpu_lane(ctx, sd, 1024, flags=_lib.WF_ACCUM64)  # another accumulation mode
    with pytest.raises(_lib.RgpuError):
        wide.resume(blob)
    # a refused snapshot leaves the lane usable once a good one is resumed
    a.resume(blob)
    a.step(1)
    # a lane with a guide attached carries network state the snapshot does not hold
    from raccoongpurender_amd import guide as G
    g = W.gpu_lane(ctx, sd, 1024)
    G.attach(g, G.Guide(ctx, seed=3))
    with pytest.raises(_lib.RgpuError):
        g.checkpoint()
    with pytest.raises(_lib.RgpuError):
        g.resume(blob)

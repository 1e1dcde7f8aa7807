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


_M = (1 << 64) - 1


def _fp_mix(h):
    h ^= h >> 30
    h = (h * 0xbf58476d1ce4e5b9) & _M
    h ^= h >> 27
    h = (h * 0x94d049bb133111eb) & _M
    return h ^ (h >> 31)


def _fingerprint64(data, seed=0):
    """The library's fingerprint of a byte string (csrc/fingerprint.h), restated for the test."""
    data = bytes(data)
    k = 0x9e3779b97f4a7c15
    lanes = [seed ^ 0x243f6a8885a308d3, seed ^ 0x13198a2e03707344, seed ^ 0xa4093822299f31d0,
             seed ^ 0x082efa98ec4e6c89]
    nb = len(data) // 32
    words = np.frombuffer(data[:nb * 32], np.uint64).reshape(nb, 4) if nb else np.zeros((0, 4), np.uint64)
    for row in words.tolist():
        for j in range(4):
            x = ((lanes[j] ^ row[j]) * k) & _M
            lanes[j] = ((x << 29) | (x >> 35)) & _M
    h = (_fp_mix(lanes[0]) ^ _fp_mix(lanes[1] + 1) ^ _fp_mix(lanes[2] + 2) ^ _fp_mix(lanes[3] + 3)
         ^ _fp_mix((len(data) + k) & _M))
    for i in range(nb * 32, len(data)):
        h = _fp_mix(h ^ data[i] ^ (i << 8))
    return _fp_mix(h)


HEADER = 208  # checkpoint header bytes; its last 8 are the fingerprint
CAM_EYE_X, CAM_WIDTH = 24, 24 + 84
PAYLOAD_FP = HEADER - 8


def _refit(blob):
    """Recompute a blob's fingerprint after an edit, as a consistent writer would."""
    b = blob.copy()
    b[PAYLOAD_FP:HEADER] = 0
    seed = _fingerprint64(b[:HEADER].tobytes())
    fp = _fingerprint64(b[HEADER:].tobytes(), seed)
    b[PAYLOAD_FP:HEADER] = np.frombuffer(np.uint64(fp).tobytes(), np.uint8)
    return b


def test_checkpoint_refuses_foreign_and_edited_snapshots(ctx):
    print("case checkpoint: foreign scene, materials, settings, edited camera and indices", flush=True)
    sd = R.cornell_box(16, 16, segments=8, rings=4)
    n = 1024
    a = W.gpu_lane(ctx, sd, n)
    a.step(3)
    blob = a.checkpoint()
    assert _refit(blob).tobytes() == blob.tobytes(), "the test's fingerprint restatement disagrees"
    before = a.state()

    def refused(lane, b, words):
        with pytest.raises(_lib.RgpuError) as ei:
            lane.resume(b)
        assert any(w in str(ei.value) for w in words), str(ei.value)

    # another scene with the same path count and image size
    other = W.gpu_lane(ctx, R.material_test_scene(16, 16), n)
    refused(other, blob, ["another scene"])
    # the same geometry with another albedo
    mats = sd["materials"]
    lam = [(Le, tuple(0.5 * r for r in R)) for Le, R in mats.lambertians]
    sd2 = dict(sd, materials=_lib.MaterialSet(mats.types, mats.indices, lam, mats.wards, mats.dielectrics,
                                              mats.volumes))
    refused(W.gpu_lane(ctx, sd2, n), blob, ["other materials"])
    # other lane settings
    refused(W.gpu_lane(ctx, sd, n, max_depth=3), blob, ["other settings"])
    # an edited camera: the fingerprint catches it; refitted, the size check does
    bad = blob.copy()
    bad[CAM_EYE_X] ^= 0x40
    refused(a, bad, ["corrupted"])
    bad = blob.copy()
    bad[CAM_WIDTH:CAM_WIDTH + 4] = np.frombuffer(np.uint32(17).tobytes(), np.uint8)
    refused(a, _refit(bad), ["camera does not match"])
    # a byte of the payload flipped
    bad = blob.copy()
    bad[HEADER + 100] ^= 1
    refused(a, bad, ["corrupted"])
    # out-of-range indices in a consistent (refitted) blob
    ray_off = HEADER
    bad = blob.copy()
    bad[ray_off + 28:ray_off + 32] = np.frombuffer(np.uint32(16 * 16).tobytes(), np.uint8)  # slot 0's pixel
    refused(a, _refit(bad), ["pixel index"])
    ext_off = HEADER + n * 64
    bad = blob.copy()
    bad[ext_off:ext_off + 4] = np.frombuffer(np.int32(1 << 20).tobytes(), np.uint8)  # slot 0's hit
    refused(a, _refit(bad), ["hit primitive"])
    items_off = HEADER + n * (32 * 5 + 16 + 1)
    counts_off = items_off + n * 7 * 4
    bad = blob.copy()
    bad[counts_off:counts_off + 4] = np.frombuffer(np.uint32(n + 1).tobytes(), np.uint8)
    refused(a, _refit(bad), ["queue is longer"])
    bad = blob.copy()
    bad[counts_off:counts_off + 4] = np.frombuffer(np.uint32(1).tobytes(), np.uint8)
    bad[items_off:items_off + 4] = np.frombuffer(np.uint32(n).tobytes(), np.uint8)
    refused(a, _refit(bad), ["queue item"])
    # every refusal came before the device was touched: the lane is as it was, and runs on
    after = a.state()
    assert W.state_mismatches(before, after) == {k: 0 for k in W.state_mismatches(before, before)}
    a.step(1)

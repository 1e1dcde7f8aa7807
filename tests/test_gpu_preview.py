"""Preview handoff between a step thread and a preview thread on two streams (the weak device mutex
of kernels/mutex.cl around accumlation_to_intermediate and copy_if_locked; the step thread and the
image thread of common/wavefront_path_tracing.hpp:1286-1313).

Bar: the lane rendered with previews running beside it ends in the state of the same render
without them (queue counts, per-pixel counts equal, radiance within 1e-4); every preview that took
the lock holds one whole published snapshot (its total sample count is the total after some
step: no torn copy), one that did not keeps its previous image, per-pixel counts in the preview
never decrease, and with the step thread idle the preview equals the lane's own intermediate bit
for bit.
"""
import threading
import time

import numpy as np
import pytest

from raccoongpurender_amd import _lib, render_scenes as R
from tests import wf_helpers as W

pytestmark = pytest.mark.gpu


def test_previews_beside_a_running_step_thread(ctx):
    print("case preview handoff: material scene 64x48, 4 paths per pixel, 80 steps", flush=True)
    sd = R.material_test_scene(64, 48, volume=True)
    P, n, steps = 64 * 48, 64 * 48 * 4, 80
    ref = W.gpu_lane(ctx, sd, n, seed=5)
    ref.step(steps)
    ref_counts, ref_cursor = ref.queue_counts()
    ref_rgb, ref_cnt = ref.accumulation()

    g = W.gpu_lane(ctx, sd, n, seed=5)
    g.set_publish(1)
    pv = ctx.stream()
    dst = ctx.alloc(P * 12)
    dst.fill_u32(0)
    took = ctx.alloc(4)
    ctx.synchronize()
    totals = {0}
    done = threading.Event()
    errors = []

    def stepper():
        try:
            for _ in range(steps):
                g.step(1)
                totals.add(g.stat())  # synchronous: the total of the snapshot this step published
        except Exception as e:  # reported by the main thread
            errors.append(e)
        finally:
            done.set()

    previews = []

    def previewer():
        try:
            while not done.is_set():
                g.update_intermediate(dst, pv, took)
                pv.synchronize()
                img = dst.download(_lib.RGB16_DTYPE, P, stream=pv)
                t = int(took.download(np.int32, 1, stream=pv)[0])
                previews.append((t, img))
                time.sleep(0.001)
        except Exception as e:
            errors.append(e)

    ts = [threading.Thread(target=stepper), threading.Thread(target=previewer)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors
    pv.synchronize()
    ctx.synchronize()

    # the render is the render without previews
    counts, cursor = g.queue_counts()
    assert np.array_equal(counts, ref_counts) and cursor == ref_cursor
    rgb, cnt = g.accumulation()
    assert np.array_equal(cnt, ref_cnt)
    assert np.allclose(rgb, ref_rgb, rtol=1e-4, atol=1e-6)

    taken = sum(t for t, _ in previews)
    print(f"{len(previews)} previews, {taken} took the lock, {len(totals) - 1} steps", flush=True)
    assert taken >= 3
    prev = np.zeros(P, _lib.RGB16_DTYPE)
    for t, img in previews:
        total = int(img["count"].sum(dtype=np.uint64))
        if t:
            assert total in totals, (total, sorted(totals)[-5:])
        else:
            assert np.array_equal(img, prev)
        assert np.all(img["count"] >= prev["count"])
        prev = img
    # nothing contends now: a preview takes the lock and equals the lane's own intermediate
    g.update_intermediate(dst, pv, took)
    pv.synchronize()
    assert int(took.download(np.int32, 1, stream=pv)[0]) == 1
    assert np.array_equal(dst.download(_lib.RGB16_DTYPE, P, stream=pv), g.intermediate())
    pv.close()


def test_graph_replays_publish_and_publishing_can_be_switched_off(ctx):
    sd = R.cornell_box(32, 32, segments=16, rings=8)
    P = 32 * 32
    g = W.gpu_lane(ctx, sd, P * 2, seed=9)
    dst = ctx.alloc(P * 12)
    took = ctx.alloc(4)
    with pytest.raises(_lib.RgpuError):
        g.update_intermediate(dst)  # publishing is off
    g.set_publish(3)
    g.step(2)  # no publish yet: the published image is the cleared one
    g.update_intermediate(dst, None, took)
    assert int(took.download(np.int32, 1)[0]) == 1
    assert int(dst.download(_lib.RGB16_DTYPE, P)["count"].sum()) == 0
    g.step(1)  # the third step publishes
    g.update_intermediate(dst)
    assert np.array_equal(dst.download(_lib.RGB16_DTYPE, P), g.intermediate())
    g.graph_capture(4)  # a replay publishes after its last step
    g.graph_launch(2)
    g.update_intermediate(dst)
    assert np.array_equal(dst.download(_lib.RGB16_DTYPE, P), g.intermediate())
    g.set_publish(0)
    with pytest.raises(_lib.RgpuError):
        g.update_intermediate(dst)

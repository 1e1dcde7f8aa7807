"""Equal-time comparison of guided and plain sampling against a high-sample reference.

relMSE = mean over pixels of (Y - Y_ref)^2 / (Y_ref^2 + 0.01), Y = luminance of the mean
radiance. All three renders run without the radiance clamp so none of them is biased.
"""
import time

import numpy as np

from . import _lib
from .guide import Guide, attach, train_pass
from .render import build_lane

NO_CLAMP = 1.0e30


def luminance_image(wf):
    rgb, cnt = wf.accumulation()
    m = rgb / np.maximum(cnt, 1)[:, None]
    return 0.2126 * m[:, 0] + 0.7152 * m[:, 1] + 0.0722 * m[:, 2], cnt


def rel_mse(y, ref):
    return float(np.mean((y - ref) ** 2 / (ref ** 2 + 1e-2)))


def rel_mse_trimmed(y, ref, drop=0.01):
    """relMSE over the pixels left after dropping the `drop` fraction with the largest error
    (caustic fireflies of the unclamped estimators)."""
    e = np.sort((y - ref) ** 2 / (ref ** 2 + 1e-2))[::-1]
    return float(e[int(e.size * drop):].mean())


def render_reference(ctx, sd, spp, paths, seed=777):
    wf = build_lane(ctx, sd, paths, clamp=NO_CLAMP, flags=_lib.WF_ACCUM64, seed=seed)
    P = sd["camera"]["width"] * sd["camera"]["height"]
    t0 = time.perf_counter()
    while True:
        wf.step(16)
        if wf.stat() >= spp * P:
            break
    y, cnt = luminance_image(wf)
    return y, time.perf_counter() - t0, float(cnt.mean())


def render_timed(ctx, sd, seconds, paths, guided, seed, train_every=4, capacity=1 << 18, select_p=0.5, lr=1e-3,
                 minibatch=16384, max_depth=0, bxdf_p=0.0, weight_cap=0.0, log_kappa_max=0.0, warmup=0.0,
                 learn_select=False):
    """`seconds` of rendering, plain or guided. A guided render trains while it renders; with
    `warmup` > 0 the image accumulated during the first `warmup` seconds (drawn while the
    network was still untrained) is discarded and accumulation restarts, training going on.
    The wall-clock budget includes the warm-up and all training."""
    wf = build_lane(ctx, sd, paths, clamp=NO_CLAMP, seed=seed, bxdf_p=bxdf_p)
    g = None
    if guided:
        lo = sd["points"].min(0) - 1e-3
        hi = sd["points"].max(0) + 1e-3
        g = Guide(ctx, lo, hi, select_p=select_p, capacity=capacity, lr=lr, seed=seed, minibatch=minibatch,
                  max_depth=max_depth, weight_cap=weight_cap, log_kappa_max=log_kappa_max, learn_select=learn_select)
        attach(wf, g, sampling=True, training=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    passes = 0
    restart = guided and warmup > 0.0
    while time.perf_counter() - t0 < seconds:
        if restart and time.perf_counter() - t0 >= warmup:
            c = sd["camera"]
            wf.initialize(_lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], c["width"], c["height"],
                                              c.get("lens_radius", 0.0)))
            restart = False
        wf.step(train_every)
        if guided:
            passes += train_pass(wf, capacity) > 0
        else:
            ctx.synchronize()
    ctx.synchronize()
    el = time.perf_counter() - t0
    y, cnt = luminance_image(wf)
    return y, el, float(cnt.mean()), passes


def equal_time_relmse(ctx, sd, seconds=4.0, ref_spp=16384, paths=1 << 20, seed=5, ref_paths=1 << 21,
                      learned_selection=True):
    """Plain and guided renders of `seconds` each (guided time includes inference and
    training) scored against a `ref_spp` plain render; with learned_selection also a guided
    render whose vertices learn their selection probability."""
    ref, t_ref, spp_ref = render_reference(ctx, sd, ref_spp, ref_paths)
    plain, t_p, spp_p, _ = render_timed(ctx, sd, seconds, paths, False, seed)
    guided, t_g, spp_g, passes = render_timed(ctx, sd, seconds, paths, True, seed)
    out = {"relmse_plain": round(rel_mse(plain, ref), 4), "relmse_guided": round(rel_mse(guided, ref), 4),
           "relmse_trim1pct_plain": round(rel_mse_trimmed(plain, ref), 4),
           "relmse_trim1pct_guided": round(rel_mse_trimmed(guided, ref), 4),
           "seconds": seconds, "plain_spp": round(spp_p, 1), "guided_spp": round(spp_g, 1),
           "guided_training_passes": passes, "reference_spp": round(spp_ref, 1), "reference_seconds": round(t_ref, 1)}
    out["ratio_plain_over_guided"] = round(out["relmse_plain"] / max(out["relmse_guided"], 1e-30), 3)
    out["ratio_trim1pct"] = round(out["relmse_trim1pct_plain"] / max(out["relmse_trim1pct_guided"], 1e-30), 3)
    if learned_selection:
        gl, _, spp_l, passes_l = render_timed(ctx, sd, seconds, paths, True, seed, learn_select=True)
        out["learned_selection"] = {
            "relmse": round(rel_mse(gl, ref), 4), "relmse_trim1pct": round(rel_mse_trimmed(gl, ref), 4),
            "spp": round(spp_l, 1), "training_passes": passes_l}
        out["learned_selection"]["ratio_plain_over_guided"] = round(
            out["relmse_plain"] / max(out["learned_selection"]["relmse"], 1e-30), 3)
        out["learned_selection"]["ratio_trim1pct"] = round(
            out["relmse_trim1pct_plain"] / max(out["learned_selection"]["relmse_trim1pct"], 1e-30), 3)
    return out


def cornell_equal_time(ctx, size=64, seconds=1.0, warmup=0.25, ref_spp=1 << 20, seed=5):
    """The Cornell box (a small ceiling light, black environment, depth 5) at size x size:
    plain against guided sampling at equal wall-clock, the guided budget including a training
    warm-up whose image is discarded; scored against a `ref_spp` plain render."""
    from .render_scenes import cornell_box

    sd = cornell_box(size, size)
    sd["max_depth"] = 5
    paths = 1 << 18
    ref, t_ref, spp_ref = render_reference(ctx, sd, ref_spp, paths)
    plain, _, spp_p, _ = render_timed(ctx, sd, seconds, paths, False, seed)
    guided, _, spp_g, passes = render_timed(ctx, sd, seconds, paths, True, seed, capacity=1 << 16, minibatch=8192,
                                            warmup=warmup)
    learned, _, spp_l, _ = render_timed(ctx, sd, seconds, paths, True, seed, capacity=1 << 16, minibatch=8192,
                                        warmup=warmup, learn_select=True)
    out = {"workload": f"Cornell box {size}x{size}, depth 5, {seconds} s each (guided: {warmup} s of training "
                       f"warm-up discarded, training continued), relMSE vs a {ref_spp}-spp plain render",
           "relmse_plain": float(f"{rel_mse(plain, ref):.4g}"), "relmse_guided": float(f"{rel_mse(guided, ref):.4g}"),
           "plain_spp": round(spp_p, 1), "guided_spp": round(spp_g, 1), "guided_training_passes": passes,
           "reference_spp": round(spp_ref, 1), "reference_seconds": round(t_ref, 1)}
    out["ratio_plain_over_guided"] = round(out["relmse_plain"] / max(out["relmse_guided"], 1e-30), 3)
    out["learned_selection"] = {"relmse": float(f"{rel_mse(learned, ref):.4g}"), "spp": round(spp_l, 1)}
    out["learned_selection"]["ratio_plain_over_guided"] = round(
        out["relmse_plain"] / max(out["learned_selection"]["relmse"], 1e-30), 3)
    return out


def main():
    """Timed renders of the full-render scene, plain then guided (the profiling entry point),
    or the whole equal-time study with --study."""
    import argparse
    import json

    from .render_workloads import full_render_scene

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--width", type=int, default=480)
    ap.add_argument("--height", type=int, default=270)
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--study", action="store_true", help="equal-time relMSE against a 16k-spp plain render")
    ap.add_argument("--paths", type=int, default=1 << 20, help="wavefront slots of the timed renders")
    ap.add_argument("--ref-paths", type=int, default=1 << 21, help="wavefront slots of the reference render")
    ap.add_argument("--ref-spp", type=int, default=16384)
    args = ap.parse_args()
    sd = full_render_scene(args.width, args.height)
    ctx = _lib.Context(0)
    if args.study:
        print(json.dumps(equal_time_relmse(ctx, sd, seconds=args.seconds, ref_spp=args.ref_spp, paths=args.paths,
                                           ref_paths=args.ref_paths)), flush=True)
    else:
        for guided in (False, True):
            _, el, spp, passes = render_timed(ctx, sd, args.seconds, 1 << 20, guided, 5)
            print(json.dumps(dict(guided=guided, seconds=round(el, 3), spp=round(spp, 1), training_passes=passes)),
                  flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

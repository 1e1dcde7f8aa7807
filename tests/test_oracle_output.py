"""The output path of the oracle (oracle.to_intermediate, merge_intermediate, tonemap_intermediate;
kernels/accumlation.cl) pinned by closed forms: half-float bit patterns of known means, zero
counts, merges that reproduce the mean of the pooled samples, and tonemap values of known means.
"""
import numpy as np

import oracle as O


def test_intermediate_known_values():
    rgb = np.array([[1.0, 2.0, 3.0], [5.0, 0.0, 0.25], [7.0, 7.0, 7.0]])
    cnt = np.array([3, 2, 0], np.uint32)
    x = O.to_intermediate(rgb, cnt)
    # 1/3, 2/3, 1 in half: 0x3555, 0x3955, 0x3c00; 2.5 = 0x4100, 0.125 = 0x3000
    assert [int(x["r"][0]), int(x["g"][0]), int(x["b"][0])] == [0x3555, 0x3955, 0x3C00]
    assert [int(x["r"][1]), int(x["g"][1]), int(x["b"][1])] == [0x4100, 0x0000, 0x3000]
    assert [int(x["r"][2]), int(x["g"][2]), int(x["b"][2]), int(x["count"][2])] == [0, 0, 0, 0]
    assert x["count"].tolist() == [3, 2, 0]


def test_merge_pools_samples():
    rng = np.random.default_rng(1)
    n = 2000
    ca = rng.integers(0, 50, n).astype(np.uint32)
    cb = rng.integers(0, 50, n).astype(np.uint32)
    sa = rng.random((n, 3)) * ca[:, None]
    sb = rng.random((n, 3)) * cb[:, None]
    m = O.merge_intermediate(O.to_intermediate(sa, ca), O.to_intermediate(sb, cb))
    pooled = O.to_intermediate(sa + sb, ca + cb)
    assert np.array_equal(m["count"], ca + cb)
    for k in "rgb":
        got = m[k].view(np.float16).astype(np.float64)
        want = pooled[k].view(np.float16).astype(np.float64)
        # two half roundings of the parts plus one of the merge: within 3 half ulps of the mean
        assert np.all(np.abs(got - want) <= 3 * 2.0 ** -10 * np.maximum(np.abs(want), 2.0 ** -14)), k
    zero = (ca + cb) == 0
    assert np.all(m["r"][zero] == 0) and np.all(m["g"][zero] == 0) and np.all(m["b"][zero] == 0)
    # merging with an empty image keeps the counts and, to one half ulp, the means
    e = O.merge_intermediate(O.to_intermediate(sa, ca), np.zeros(n, O.RGB16_DTYPE))
    base = O.to_intermediate(sa, ca)
    assert np.array_equal(e["count"], ca)
    assert np.all(np.abs(e["r"].astype(np.int64) - base["r"].astype(np.int64)) <= 1)


def test_tonemap_known_values():
    x = O.to_intermediate(np.array([[1.0, 0.0, 0.5], [100.0, 0.25, 0.0]]), np.array([1, 1], np.uint32))
    img = O.tonemap_intermediate(x)
    # pow(1, 1/2.2) * 256 = 256 -> 255; 0 -> 0; 0.5^(1/2.2) * 256 = 186.8 -> 186; 0.25^(1/2.2) * 256 = 136.3
    assert img[0].tolist() == [255, 0, 186, 255]
    assert img[1].tolist() == [255, 136, 0, 255]

"""Compressed stackless records (RGPU_VARIANT_COMPRESSED) against the exact tree they encode.

The packer must round every child plane outward: a decoded box, computed in float exactly as
the kernels compute it (origin + q * 2^(e - 127), the product exact, one rounding in the sum),
contains the child's exact box, so a ray that the exact walk sends into a child enters the
decoded child as well. The layout (root at record 0, children in pairs, leaf ranges, up words
with the sibling's leaf bit) is checked by walking the built tree and the records together.
"""
import numpy as np
import pytest

from raccoongpurender_amd import _lib
from tests.test_bvh_traverse_cpu import _meshes

LEAF = np.uint32(0x80000000)


def _decode(rec, r):
    f = rec[r].view(np.float32)
    u = rec[r].view(np.uint32)
    q = rec[r, 16:28]
    exps = int(u[3])
    planes = np.zeros((4, 3), np.float32)  # Lmin, Lmax, Rmin, Rmax
    for k in range(3):
        scale = np.array([((exps >> (8 * k)) & 0xFF) << 23], np.uint32).view(np.float32)[0]
        for j in range(4):
            planes[j, k] = np.float32(f[k] + np.float32(np.float32(q[3 * j + k]) * scale))
    return planes, exps >> 24, int(u[7])


@pytest.mark.parametrize("max_leaf", [1, 4])
def test_compressed_records_contain_the_exact_boxes(max_leaf):
    for name, v, f, sph in _meshes():
        bvh = _lib.BVH(v, f, sph, max_leaf=max_leaf)
        t = bvh.export()
        rec, up = bvh.export_compressed()
        assert len(rec) == bvh.nodes + 1 and up[0] == 0xFFFFFFFF
        b = t["bounds"]
        used = np.zeros(len(rec), int)
        loose = []
        stack = [(0, 0)]
        while stack:
            i, r = stack.pop()
            used[r] += 1
            if t["left"][i] < 0:
                u = rec[r].view(np.uint32)
                assert u[0] == t["prim_beg"][i] and u[1] == t["prim_beg"][i] + t["prim_cnt"][i], name
                continue
            L, R = t["left"][i], t["right"][i]
            planes, flags, child = _decode(rec, r)
            assert child % 2 == 0 and child >= 2, name
            assert flags == (1 if t["left"][L] < 0 else 0) | (2 if t["left"][R] < 0 else 0), name
            for c, (lo, hi) in ((L, (0, 1)), (R, (2, 3))):
                assert np.all(planes[lo] <= b[c, :3]) and np.all(planes[hi] >= b[c, 3:]), (name, i)
                ext = np.maximum(b[i, 3:] - b[i, :3], 1e-30)
                loose.append(((b[c, :3] - planes[lo]) / ext).max())
                loose.append(((planes[hi] - b[c, 3:]) / ext).max())
            assert up[child] == r | (LEAF if t["left"][R] < 0 else 0), name
            assert up[child + 1] == r | (LEAF if t["left"][L] < 0 else 0), name
            stack += [(L, child), (R, child + 1)]
        assert used[0] == 1 and used[1] == 0 and np.all(used[2:] == 1), name
        # outward rounding costs at most a couple of grid steps (1/255 of the parent's extent)
        print(f"{name}: mean loosening {np.mean(loose):.5f}, max {np.max(loose):.5f} of the parent extent")
        assert np.max(loose) <= 2.0 / 255.0 + 1e-6, name


def test_single_leaf_tree_is_one_record():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    bvh = _lib.BVH(v, f)
    rec, up = bvh.export_compressed()
    assert len(rec) == 2 and up[0] == 0xFFFFFFFF
    assert rec[0].view(np.uint32)[0] == 0 and rec[0].view(np.uint32)[1] == 1

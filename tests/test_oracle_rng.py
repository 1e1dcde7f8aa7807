"""Pins the oracle's RNG (kernels/peseudo_random.cl:5-64, common/share/peseudo_random.hpp:32-142)."""
import json
import os

import numpy as np

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_kat.json")))


def test_splitmix64_known_answers():
    got = O.splitmix64_seq(0, 4)
    assert [int(x) for x in got] == [int(h, 16) for h in GOLD["splitmix64_seed0_first4"]]


def test_xoshiro128ss_known_answers():
    s = np.array([1, 2, 3, 4], np.uint32)
    assert [int(x) for x in O.xoshiro_next(s, 10)] == GOLD["xoshiro128ss_state_1_2_3_4_first10"]


def _transition_matrix():
    """128x128 GF(2) matrix of one xoshiro128** state step, from the oracle's own step."""
    T = np.zeros((128, 128), np.uint8)
    for j in range(128):
        s = np.zeros(4, np.uint32)
        s[j // 32] = np.uint32(1 << (j % 32))
        O.xoshiro_next(s, 1)
        bits = np.unpackbits(s.view(np.uint8), bitorder="little")
        T[:, j] = bits
    return T


def _gf2_mul(A, B):
    return (A.astype(np.int64) @ B.astype(np.int64) % 2).astype(np.uint8)


def test_xoshiro_engine_matches_published_jump_polynomial():
    # The published jump constants encode x^(2^64) modulo the characteristic polynomial
    # of the correct linear engine. Squaring the oracle's transition matrix 64 times
    # must therefore reproduce the jump function; a wrong shift or rotation breaks it.
    T = _transition_matrix()
    M = T.copy()
    for _ in range(64):
        M = _gf2_mul(M, M)
    rng = np.random.default_rng(3)
    for _ in range(4):
        s = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        bits = np.unpackbits(s.view(np.uint8), bitorder="little")
        expect = np.packbits(_gf2_mul(M, bits[:, None])[:, 0], bitorder="little").view(np.uint32)
        j = s.copy()
        O.xoshiro_jump(j)
        assert np.array_equal(j, expect)


def test_random_initialize_matches_splitmix_split():
    # kernels/peseudo_random.cl:48-64: state = (lo(r0), hi(r0), lo(r1), hi(r1)) of splitmix64(seed_offset+gid)
    st = O.random_initialize(1000, 5)
    for gid in range(5):
        r = O.splitmix64_seq(1000 + gid, 2)
        expect = [int(r[0]) & 0xFFFFFFFF, int(r[0]) >> 32, int(r[1]) & 0xFFFFFFFF, int(r[1]) >> 32]
        assert [int(x) for x in st[gid]] == expect


def test_uniform_float_and_integer_ranges():
    s = O.random_initialize(7, 1)[0].copy()
    u = O.xoshiro_uniform(s, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005
    # 23-bit mantissa grid: every value is k * 2^-23
    assert np.all((u * np.float32(2**23)) == np.floor(u * np.float32(2**23)))
    s2 = O.random_initialize(7, 1)[0].copy()
    raw = O.xoshiro_next(s2, 2)
    s3 = O.random_initialize(7, 1)[0].copy()
    w = O.xoshiro_uniform_integer(s3, 1)[0]
    assert int(w) == ((int(raw[0]) >> 1) << 31) | (int(raw[1]) >> 1)
    assert int(w) < 2**62
    # uniform = (next >> 9 | 0x3f800000) - 1
    s4 = O.random_initialize(9, 1)[0].copy()
    r4 = O.xoshiro_next(s4.copy(), 3)
    f4 = O.xoshiro_uniform(s4, 3)
    expect = ((r4 >> 9) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
    assert np.array_equal(f4, expect)

"""Pins the CPU oracle: hand-computed known-answer vectors, closed forms and
independent algorithms (no GPU).  These are the checks that make the oracle
trustworthy as the parity reference for the HIP kernels."""
import numpy as np
import pytest

UNIT_MIN = (0.0, 0.0, 0.0)
UNIT_MAX = (1.0, 1.0, 1.0)


def ray(o, d, tmin=0.0, tmax=1e30):
    return [o[0], o[1], o[2], tmin, d[0], d[1], d[2], tmax]


# --------------------------------------------------------------- Aila32
# Each literal was derived by hand from the bit list of
# computeMortonCodesAila32.comp:47-88:
#  origin 0, dir +x : ia=(0,0,0), ib=(15,7,7) -> bits 19, 13,12,11, 7,6,5, 1,0
#  origin 1, dir -z : ia=(127,127,127), ib=(7,7,0) -> origin bits all set,
#                     direction bits 13,12, 7,6, 1,0
AILA_CASES = [
    (ray((0, 0, 0), (1, 0, 0)), 0x000838E3),
    (ray((1, 1, 1), (0, 0, -1)), 0xFFF1F7DF),
    # direction scale does not matter (normalised): same as case 1
    (ray((0, 0, 0), (5, 0, 0)), 0x000838E3),
    # origin below the box clamps to 0, dir -x: ib=(0,7,7) -> bits 12,11, 6,5, 0
    (ray((-3, -3, -3), (-1, 0, 0)), (1 << 12) | (1 << 11) | (1 << 6) | (1 << 5) | 1),
]


@pytest.mark.parametrize("r,expected", AILA_CASES)
def test_aila32_known_answers(orc, r, expected):
    got = orc.keys_aila32(np.array([r], np.float32), UNIT_MIN, UNIT_MAX)
    assert int(got[0]) == expected, hex(int(got[0]))


def test_aila32_scene_frame(orc):
    # the same ray expressed in a translated and scaled scene gives the same key
    r = np.array([ray((0.5, 0.25, 0.75), (0.3, -0.4, 0.5))], np.float32)
    k0 = orc.keys_aila32(r, UNIT_MIN, UNIT_MAX)
    r2 = r.copy()
    r2[:, :3] = r2[:, :3] * 4.0 + 2.0
    k1 = orc.keys_aila32(r2, (2, 2, 2), (6, 6, 6))
    assert k0[0] == k1[0]


def test_aila32_origin_bits_order(orc):
    # Moving the origin along x only changes bits assigned to x.
    xs = np.linspace(0, 1, 128, dtype=np.float32)
    rays = np.array([ray((x, 0, 0), (0, 0, 1)) for x in xs], np.float32)
    keys = orc.keys_aila32(rays, UNIT_MIN, UNIT_MAX)
    x_bits = (1 << 31) | (1 << 28) | (1 << 25) | (1 << 22) | (1 << 16) | (1 << 10) | (1 << 4)
    base = int(keys[0])
    for k in keys:
        assert (int(k) ^ base) & ~x_bits == 0
    # and the x field is monotone in x
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


# -------------------------------------------------------- 30 / 60 bit keys
def test_origin30_known_answers(orc):
    rays = np.array([ray((0, 0, 0), (1, 0, 0)), ray((1, 1, 1), (1, 0, 0)),
                     ray((1, 0, 0), (1, 0, 0)), ray((0, 0, 1), (1, 0, 0))], np.float32)
    k = orc.keys_origin30(rays, UNIT_MIN, UNIT_MAX)
    assert [int(x) for x in k] == [0, 0x3FFFFFFF, 0x24924924, 0x09249249]


def test_oct60_known_answers(orc):
    rays = np.array([ray((0, 0, 0), (0, 0, 1)),     # u=v=0.5 -> q=16383
                     ray((0, 0, 0), (0, 0, -1)),    # folded to u=v=1 -> q=32767
                     ray((0, 0, 0), (-1, 0, 0)),    # u=0, v=0.5
                     ray((1, 1, 1), (0, 0, 1))], np.float32)
    k = orc.keys_oct60(rays, UNIT_MIN, UNIT_MAX)
    assert int(k[0]) == 0x0FFFFFFF
    assert int(k[1]) == 0x3FFFFFFF
    assert int(k[2]) == 0x05555555
    assert int(k[3]) == (0x3FFFFFFF << 30) | 0x0FFFFFFF


def test_scene64_prefix(orc):
    rays = np.array([ray((0, 0, 0), (0, 0, 1))] * 3, np.float32)
    scene = np.array([0, 5, 63], np.uint8)
    mins = np.zeros((64, 3), np.float32)
    maxs = np.ones((64, 3), np.float32)
    k = orc.keys_scene64(rays, scene, mins, maxs)
    # 14-bit octahedral code of +z: q = uint(0.5*16383) = 8191 -> 26 low bits set
    low = (1 << 26) - 1
    assert [int(x) for x in k] == [low, (5 << 58) | low, (63 << 58) | low]


# ------------------------------------------------------------------- sort
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 32), (np.uint32, 30), (np.uint64, 60),
                                        (np.uint64, 64), (np.uint32, 7)])
@pytest.mark.parametrize("n", [0, 1, 2, 17, 1000, 4097])
def test_sort_matches_numpy_stable(orc, dtype, bits, n):
    rng = np.random.default_rng(n + bits)
    hi = np.iinfo(dtype).max
    keys = rng.integers(0, hi, size=n, dtype=dtype, endpoint=True)
    keys[: n // 3] = keys[0] if n else keys[: n // 3]   # many duplicates
    vals = np.arange(n, dtype=np.uint32)
    k, v = orc.sort_pairs(keys, vals, bits)
    mask = (1 << bits) - 1
    masked = (keys.astype(object) & mask).astype(np.uint64) if n else keys.astype(np.uint64)
    order = np.argsort(masked, kind="stable")
    assert np.array_equal(v, order.astype(np.uint32))
    assert np.array_equal(k, keys[order])


# ---------------------------------------------------------------- reorder
def test_reorder_invariants(orc):
    rng = np.random.default_rng(3)
    n = 5000
    rays = rng.standard_normal((n, 8)).astype(np.float32)
    perm = rng.permutation(n).astype(np.uint32)
    s2i = rng.permutation(2 * n)[:n].astype(np.int32)   # pixel ids, sparse
    out_r, out_s2i, out_i2s = orc.reorder(perm, rays, s2i, 2 * n)
    assert np.array_equal(out_r, rays[perm])
    assert np.array_equal(out_s2i, s2i[perm])
    # i2s inverts s2i on the written entries
    assert np.array_equal(out_i2s[out_s2i], np.arange(n, dtype=np.int32))
    assert np.sum(out_i2s >= 0) == n

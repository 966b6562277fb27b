"""The eta bands of the duplicate-removal kernels (stages_clean.hip) visit a
superset of the pairs that can flag.

The T5 (SDL/Kernels.h:250-320) and pLS (SDL/Kernels.h:384-473) rules skip a
pair when float |eta_a - eta_b| > 0.1, so only pairs with |d eta| <= 0.1 or a
NaN d eta can flag.  The kernels bin eta into 1/8-wide bins (ends clamped, NaN
apart) and let a tile of x meet bins bmin-1 .. bmax+1 plus the NaN bin.  This
checks, in float32 as on the device, that every pair that can flag lands in
bins at most one apart or involves NaN.
"""
import numpy as np

K_BINS = 64
NAN_BIN = K_BINS


def eta_bin(eta):
    eta = np.float32(eta)
    if eta != eta:
        return NAN_BIN
    if not eta > np.float32(-4):
        return 0
    if not eta < np.float32(4):
        return K_BINS - 1
    b = int(np.float32(np.float32(eta + np.float32(4)) * np.float32(8)))
    return b if b < K_BINS else K_BINS - 1


def can_flag(a, b):
    d = np.float32(abs(np.float32(np.float32(a) - np.float32(b))))
    return not (d > np.float32(0.1))


def visited(a, b):
    ba, bb = eta_bin(a), eta_bin(b)
    return ba == NAN_BIN or bb == NAN_BIN or abs(ba - bb) <= 1


def test_pairs_that_can_flag_are_visited_near_bin_edges():
    rng = np.random.default_rng(1)
    edges = np.arange(-4.5, 4.5, 0.125, dtype=np.float64)
    checked = 0
    for e in edges:
        for da in np.concatenate([rng.uniform(-2e-6, 2e-6, 20), [0.0]]):
            a = np.float32(e + da)
            for db in (0.1, 0.0999999, 0.1000001, -0.1, -0.0999999, 0.05, 0.0):
                b = np.float32(a + np.float32(db))
                for lo in (np.nextafter(b, np.float32(-np.inf)), b, np.nextafter(b, np.float32(np.inf))):
                    if can_flag(a, lo):
                        assert visited(a, lo), (a, lo)
                        checked += 1
    assert checked > 1000


def test_random_pairs_and_special_values():
    rng = np.random.default_rng(2)
    x = rng.uniform(-5, 5, 4000).astype(np.float32)
    y = (x + rng.uniform(-0.11, 0.11, 4000)).astype(np.float32)
    for a, b in zip(x, y):
        if can_flag(a, b):
            assert visited(a, b)
    inf = np.float32(np.inf)
    for a, b in [(np.nan, 0.3), (inf, inf), (-inf, -inf), (inf, 3.99), (-4.0, -3.95), (3.95, 4.02)]:
        if can_flag(a, b):
            assert visited(a, b), (a, b)

"""Integer steps of the device kernels that use float arithmetic, checked on
the CPU with the same IEEE float32 operations (numpy float32 rounds each
operation to nearest, as the device code compiled with correctly rounded
division does)."""
import numpy as np


def small_div(n: np.ndarray, d: int) -> np.ndarray:
    """The tile index of a cell in the deposition's list passes (small_div in
    asmooth_amd/csrc/deposit.hip): the quotient from the float reciprocal,
    truncated, then one correction step."""
    inv = np.float32(1.0) / np.float32(d)
    q = (n.astype(np.float32) * inv).astype(np.int64)
    q = np.where(q * d > n, q - 1, np.where((q + 1) * d <= n, q + 1, q))
    return q


def test_small_div_is_the_integer_quotient():
    # every cell index a mesh of up to 2^22 cells per side can hold, for
    # every tile side the deposition uses (1..8) and a few beyond
    n = np.arange(1 << 22, dtype=np.int64)
    for d in list(range(1, 9)) + [11, 16, 31, 64, 255, 256]:
        assert np.array_equal(small_div(n, d), n // d), d


def test_small_div_needs_its_correction():
    # Tile sides 1..8 never need it below 2^22, but the truncated product is
    # one short of exact multiples for some divisors (41 at n = 41: the
    # reciprocal rounds down), so the step is what makes the helper exact in
    # general; it must fix every such case.
    n = np.arange(1 << 22, dtype=np.int64)
    inv = np.float32(1.0) / np.float32(41)
    q = (n.astype(np.float32) * inv).astype(np.int64)
    assert np.count_nonzero(q != n // 41) > 0
    assert np.array_equal(small_div(n, 41), n // 41)
    for d in (47, 61, 97, 109, 127, 200):
        assert np.array_equal(small_div(n, d), n // d), d

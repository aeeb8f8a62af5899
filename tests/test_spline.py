"""The cubic spline option (no GPU).

Density kernel: the oracle's orc_spline against the reference's own
spline_ker (nbr_search.cpp:39-53, compiled from its source in
oracle/_ref/libasmref_spline.so) bit for bit, and against closed forms:
W(0) = 8 / (pi h^3), continuity at r = 1/2, zero at r = 1, unit integral.
Mesh kernel (no counterpart in the reference, which deposits only the tent):
the separable cubic B-spline's cell integral against quadrature of its
definition, partition of unity and symmetry, and a one-particle deposition
equal to the outer product of three such integrals.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O


def _w1(t):
    """The 1-D cubic B-spline of half-width 1, normalised to one."""
    t = abs(t)
    if t < 0.5:
        return 4.0 / 3.0 * (1.0 - 6.0 * t * t + 6.0 * t ** 3)
    if t <= 1.0:
        return 8.0 / 3.0 * (1.0 - t) ** 3
    return 0.0


@pytest.mark.skipif(not O.have_ref_spline(), reason="oracle/_ref not built (reference tree absent)")
def test_spline_density_kernel_is_the_references_bit_for_bit():
    rng = np.random.default_rng(3)
    for _ in range(20000):
        h = np.float32(rng.uniform(0.05, 20.0))
        hinv = np.float32(1.0) / h
        x = rng.uniform(-50, 50, 3).astype(np.float32)
        u = rng.normal(size=3)
        r = rng.choice([rng.uniform(0, 1), rng.uniform(0.49, 0.51), rng.uniform(0.99, 1.0), 0.0])
        y = (x + (u / np.linalg.norm(u) * r * float(h))).astype(np.float32)
        if math.sqrt(float(O.norm2(x, y))) * float(hinv) > 1.00001:   # outside the reference's assert
            continue
        got = np.float32(O.spline(x, y, hinv))
        ref = np.float32(O.ref_spline_ker(x, y, hinv))
        assert got.view(np.int32) == ref.view(np.int32), (x, y, hinv, got, ref)


def test_spline_density_kernel_closed_forms():
    h = 2.5
    hinv = np.float32(1.0 / h)
    x = np.zeros(3, np.float32)
    w0 = O.spline(x, x, hinv)
    assert w0 == np.float32(8.0 / math.pi * float(hinv) ** 3)
    # continuity at r = 1/2 and zero at r = 1
    f = lambda r: O.spline(x, np.array([r * h, 0, 0], np.float32), hinv)
    assert f(0.5 * (1 - 1e-6)) == pytest.approx(f(0.5 * (1 + 1e-6)), rel=1e-4)
    assert f(0.5 * (1 + 1e-6)) == pytest.approx(w0 * 0.25, rel=1e-4)
    assert abs(f(1.0)) < 1e-7 * w0
    # integral over all space: 4 pi int_0^h W(r) r^2 dr = 1
    g = lambda d: O.spline(x, np.array([d, 0, 0], np.float32), hinv)   # at distance d
    total, _ = integrate.quad(lambda d: 4 * math.pi * d * d * g(d), 0.0, h, points=[0.5 * h], limit=200)
    assert total == pytest.approx(1.0, abs=2e-6)


def test_spline_cell_integral_against_quadrature():
    rng = np.random.default_rng(4)
    for _ in range(300):
        a, b = sorted(rng.uniform(-1.3, 1.3, 2))
        got = O.spline_integ(a, b)
        want, _ = integrate.quad(_w1, a, b, points=[-1, -0.5, 0, 0.5, 1], limit=200)
        assert got == pytest.approx(want, abs=3e-7)
        # symmetry: the kernel is even
        assert O.spline_integ(-b, -a) == pytest.approx(got, abs=3e-7)
    assert O.spline_integ(-1.0, 1.0) == pytest.approx(1.0, abs=1e-7)
    # partition of unity over cells of any size
    for ds in (0.07, 0.3, 0.8):
        edges = np.arange(-1.5, 1.5 + ds, ds, dtype=np.float32)
        s = sum(O.spline_integ(edges[i], edges[i] + np.float32(ds)) for i in range(len(edges)))
        assert s == pytest.approx(1.0, abs=2e-6)


def test_spline_single_particle_deposition_is_an_outer_product():
    L, nc = 16.0, 16
    x = np.array([[7.3, 8.1, 8.6]], np.float32)
    h = np.array([3.1], np.float32)
    mesh = O.assign_on_mesh(x, np.ones((1, 3), np.float32), np.ones(1, np.float32), h, np.zeros(1, np.int32),
                            L, nc, -1, kern=1)
    m = nc ** 3
    n = mesh[:m].reshape(nc, nc, nc)
    assert n.sum() == pytest.approx(1.0, abs=1e-6)
    dx = L / nc
    w = [np.array([integrate.quad(lambda t: _w1(t), (i * dx - x[0, a]) / h[0], ((i + 1) * dx - x[0, a]) / h[0])[0]
                   for i in range(nc)]) for a in range(3)]
    want = w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    assert np.max(np.abs(n - want)) < 2e-6 * want.max()
    # the tent gives a different, also normalised, weighting
    tent = O.assign_on_mesh(x, np.ones((1, 3), np.float32), np.ones(1, np.float32), h, np.zeros(1, np.int32),
                            L, nc, -1)
    assert tent[:m].sum() == pytest.approx(1.0, abs=1e-6)
    assert np.max(np.abs(tent[:m] - mesh[:m])) > 1e-3

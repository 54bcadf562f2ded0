"""Oracle pins: the reference's four manufactured-solution studies.

The reference prints convergence tables (tests/Poisson_test.cpp,
IMEX_LDG_test.cpp, DD_Poisson_test.cpp, interface_test.cpp) and asserts
nothing; the textbook rates are asserted here: LDG density O(h^{k+1}),
LDG current O(h^k), mixed Phi O(h^k), mixed D at least O(h^k).
"""
import numpy as np
import pytest

from oracle import manufactured as mm


def _rates(rows, key):
    return [np.log2(rows[i - 1][key] / rows[i][key]) for i in range(1, len(rows))]


@pytest.mark.parametrize("k", [1, 2])
def test_steady_state_rates(k):
    rows = [mm.steady_state(k, r) for r in range(2, 5)]
    r_phi = _rates(rows, "Phi")[-1]
    r_d = _rates(rows, "D")[-1]
    r_u = _rates(rows, "u")[-1]
    r_j = _rates(rows, "J")[-1]
    assert r_phi > k - 0.15, r_phi
    assert r_d > k - 0.15, r_d
    assert r_u > k + 1 - 0.15, r_u
    assert r_j > k - 0.15, r_j


def test_transient_rates():
    rows = [mm.transient(1, r) for r in range(2, 5)]
    assert _rates(rows, "u")[-1] > 1.85
    assert _rates(rows, "J")[-1] > 0.9


def test_transient_reference_ic_quirk():
    # the reference projects 1 + cos(2 pi x) cos(2 pi y) instead of u(0)
    # (test_functions.cpp:429-448); the error is then slightly larger
    a = mm.transient(1, 3, ic="exact")["u"]
    b = mm.transient(1, 3, ic="reference")["u"]
    assert b > a


def test_dd_poisson_rates():
    rows = [mm.dd_poisson(1, r) for r in range(2, 5)]
    assert _rates(rows, "u")[-1] > 1.8
    assert _rates(rows, "Phi")[-1] > 0.9
    assert _rates(rows, "D")[-1] > 0.9


def test_interface_rates():
    rows = [mm.interface(2, r) for r in range(1, 4)]
    assert _rates(rows, "u1")[-1] > 2.8
    assert _rates(rows, "u2")[-1] > 2.8


# The same studies on the solar-cell grid the GPU path runs them on
# (geometry='cell': semiconductor half with the exact flux on x = 0.5, the
# Poisson problem on the merged square with Dirichlet data all round).
@pytest.mark.parametrize("k", [1, 2])
def test_steady_state_rates_cell_grid(k):
    rows = [mm.steady_state(k, r, geometry="cell") for r in range(2, 5)]
    assert _rates(rows, "Phi")[-1] > k - 0.15
    assert _rates(rows, "D")[-1] > k + 1 - 0.15
    assert _rates(rows, "u")[-1] > k + 1 - 0.15
    assert _rates(rows, "J")[-1] > k - 0.15


def test_transient_and_dd_rates_cell_grid():
    rows = [mm.transient(1, r, geometry="cell") for r in range(2, 5)]
    assert _rates(rows, "u")[-1] > 1.85
    assert _rates(rows, "J")[-1] > 0.8
    rows = [mm.dd_poisson(1, r, geometry="cell") for r in range(2, 5)]
    assert _rates(rows, "u")[-1] > 1.8
    assert _rates(rows, "Phi")[-1] > 0.9
    assert _rates(rows, "D")[-1] > 1.8


def test_exact_fluxes():
    # the reference's transient interface datum (x = 1) is the exact J.n there,
    # and the flux the 'cell' grid prescribes on x = 0.5 is J.n of the same
    # exact solutions: the currents are minus the gradients (finite differences)
    rng = np.random.default_rng(3)
    y = rng.uniform(0, 1, 16)
    t = 0.37
    x1 = np.stack([np.ones_like(y), y], -1)
    ref = -np.exp(-t) - np.cos(2 * np.pi * y) - 1.0
    assert np.abs(mm.transient_J(x1, t)[..., 0] - ref).max() < 1e-14
    eps = 1e-6
    pts = np.stack([rng.uniform(0, 1, 16), y], -1)
    ex, ey = np.array([eps, 0.0]), np.array([0.0, eps])
    du = np.stack([(mm.u_exact(pts + ex, t) - mm.u_exact(pts - ex, t)) / (2 * eps),
                   (mm.u_exact(pts + ey, t) - mm.u_exact(pts - ey, t)) / (2 * eps)], -1)
    dg = np.stack([(mm.poisson_g(pts + ex) - mm.poisson_g(pts - ex)) / (2 * eps),
                   (mm.poisson_g(pts + ey) - mm.poisson_g(pts - ey)) / (2 * eps)], -1)
    u = mm.u_exact(pts, t)[..., None]
    assert np.abs(mm.poisson_D(pts) + dg).max() < 1e-7                              # D = -grad Phi
    assert np.abs(mm.transient_J(pts, t) - (-du - np.array([1.0, 0.0]) * u)).max() < 1e-7
    assert np.abs(mm.dd_J(pts, t) - (-du - mm.poisson_D(pts) * u)).max() < 1e-7

"""The oracle's constant matrices against exact sympy fixtures (tests/golden/).

ldg_*.json: the LDG system matrix A_1 and the mass matrix M on a 2 x 2
rectangular region, restated from LDG.cpp:40-678 and integrated exactly.
mixed_*.json: the RT-basis-independent potential Schur complement C A^{-1} B
of the mixed Poisson matrix (MixedFEM.cpp:29-164) with D.n = 0 on Neumann
faces.  On rectangles Gauss quadrature is exact, so agreement is to rounding.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.ldg import CarrierFE, assemble_ldg
from oracle.mesh import SolarCellMesh
from oracle.mixed import PoissonFE, assemble_mixed, constrained_dofs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GEO = dict(r1=0.5, r2=0.5, H=1.0, L=1.0)


def _load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def _sparse(d):
    n = d["n"]
    return sp.coo_matrix((d["v"], (d["i"], d["j"])), shape=(n, n)).toarray()


@pytest.mark.parametrize("case", ["k1_r1", "k2_r1_insulated"])
def test_ldg_matrix_matches_exact(case):
    fx = _load(f"ldg_{case}.json")
    p = fx["params"]
    k, r = int(p["k"]), int(p["r"])
    mesh = SolarCellMesh(r, GEO, bool(p["insulated"]), False)
    cfe = CarrierFE(k)
    A1, A2, M = assemble_ldg(mesh.semi, cfe, p["mu"], 2.0 * p["mu"], p["dt"], p["theta"], p["pen"])
    Aref, Mref = _sparse(fx["A"]), _sparse(fx["M"])
    assert A1.shape == Aref.shape
    scale = np.abs(Aref).max()
    assert np.abs(A1.toarray() - Aref).max() <= 1e-13 * scale
    assert np.abs(M.toarray() - Mref).max() <= 1e-13 * np.abs(Mref).max()
    # the second carrier differs only in the current mass block (1/mu)
    assert np.abs((A1 - A2).toarray()).max() > 0


def test_ldg_fixture_detects_planted_mistakes():
    """A flipped beta sign or a dropped penalty in the oracle's flux must fail."""
    fx = _load("ldg_k1_r1.json")
    p = fx["params"]
    mesh = SolarCellMesh(1, GEO, False, False)
    cfe = CarrierFE(1)
    Aref = _sparse(fx["A"])
    import oracle.ldg as L
    saved = L.BETA.copy()
    try:
        L.BETA[:] = -saved
        A1, _, _ = assemble_ldg(mesh.semi, cfe, p["mu"], p["mu"], p["dt"], p["theta"], p["pen"])
        assert np.abs(A1.toarray() - Aref).max() > 1e-3
    finally:
        L.BETA[:] = saved
    A1, _, _ = assemble_ldg(mesh.semi, cfe, p["mu"], p["mu"], p["dt"], p["theta"], 0.0)
    assert np.abs(A1.toarray() - Aref).max() > 1e-3


@pytest.mark.parametrize("case", ["k1_r1", "k2_r1_insulated"])
def test_mixed_schur_matches_exact(case):
    fx = _load(f"mixed_{case}.json")
    p = fx["params"]
    k, r = int(p["k"]), int(p["r"])
    mesh = SolarCellMesh(r, GEO, bool(p["insulated"]), False)
    pfe = PoissonFE(k, mesh)
    A = assemble_mixed(mesh, pfe, p["eps_s"], p["eps_e"], p["lam2"]).toarray()
    cons = set(constrained_dofs(mesh, pfe).tolist())
    nphi = mesh.ncell_poisson * k * k if hasattr(mesh, "ncell_poisson") else None
    Sref = np.array(fx["S"])
    nphi = Sref.shape[0]
    nD = A.shape[0] - nphi
    freeD = [i for i in range(nD) if i not in cons]
    AA = A[np.ix_(freeD, freeD)]
    B = A[np.ix_(freeD, range(nD, nD + nphi))]
    C = A[np.ix_(range(nD, nD + nphi), freeD)]
    S = C @ np.linalg.solve(AA, B)
    assert np.abs(S - Sref).max() <= 1e-12 * np.abs(Sref).max()

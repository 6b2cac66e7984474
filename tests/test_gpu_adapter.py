"""The reference edge-kernel unit tests run through the C++ adapter
(tests/cpp/unit_edge_kernels.cpp over nalu-wind_amd/host/nalu_adapter.h):
the driver assembles each fixture on the GPU and this test compares the
systems it prints with the reference golds."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nalu-wind_amd", "build", "unit_edge_kernels")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_golds.json")))


@pytest.fixture(scope="module")
def out():
    if not os.path.exists(EXE):
        pytest.fail("adapter test driver not built (nalu-wind_amd/build.py)")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _dense(sysd):
    rp, cols, vals = sysd["rowOffsets"], sysd["cols"], sysd["vals"]
    n = len(rp) - 1
    D = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            D[r, cols[k]] += vals[k]
    return D


def test_continuity_advection(out):
    g = GOLD["continuity_adv_edge"]
    s = out["continuity_advection"]
    np.testing.assert_allclose(s["rhs"], g["rhs"], rtol=0, atol=g["rhs_tol"])
    np.testing.assert_allclose(_dense(s).ravel(), g["lhs"], rtol=0, atol=g["lhs_tol"])


def test_momentum_advection_diffusion(out):
    g = GOLD["momentum_adv_diff_edge"]
    s = out["momentum_advection_diffusion"]
    D = _dense(s)
    G = np.array(g["lhs"]).reshape(24, 24)
    same = np.zeros((24, 24), bool)
    for d in range(3):
        same[d::3, d::3] = True
    np.testing.assert_allclose(s["rhs"], g["rhs"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(D[same], G[same], rtol=0, atol=g["tol"])


@pytest.mark.parametrize("name,vals,rhs", [
    ("scalar_adv_diff", "vals_serial", "rhs_serial"),
    ("scalar_adv_diff_fix_pressure", "fixed_vals_serial", "fixed_rhs_serial"),
    ("scalar_adv_diff_dirichlet", "dirichlet_vals_serial", "dirichlet_rhs_serial")])
def test_scalar_adv_diff_tpetra(out, name, vals, rhs):
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    s = out[name]
    assert s["rowOffsets"] == g["rowOffsets_serial"]
    assert s["cols"] == g["cols_serial"]
    np.testing.assert_allclose(s["vals"], g[vals], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(s["rhs"], g[rhs], rtol=0, atol=g["tol"])


def test_mdot_calc_edge(out):
    """UnitTestMdotAlg.C:21-78 through MdotEdgeAlg of the adapter."""
    g = GOLD["mdot_edge_unit_hex"]
    md = np.array(out["mdot_calc_edge"])
    assert md.size == g["num_edges"]
    np.testing.assert_allclose(md, g["mdot"], rtol=0, atol=g["tol"])


def test_geometry_unit_hex(out):
    """UnitTestGeometryAlg.C:20-100 through GeometryInteriorAlg of the
    adapter: dual volume 0.125 per node, |A|^2 = 0.0625 per edge, and the
    same area vectors as the host producer."""
    g = GOLD["geometry_unit_hex"]
    r = out["geometry_unit_hex"]
    np.testing.assert_allclose(r["dual_volume"], g["dual_volume"], rtol=0, atol=g["tol"])
    a = np.array(r["area_vector"]).reshape(-1, 3)
    assert a.shape[0] == g["num_edges"]
    np.testing.assert_allclose((a ** 2).sum(1), g["area_mag_sq"], rtol=0, atol=g["tol"])

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


def _got(sysd):
    return (np.array(sysd["rowOffsets"]), np.array(sysd["cols"]),
            np.array(sysd["vals"]), np.array(sysd["rhs"]))


def _oracle_two_systems(t):
    from oracle import orc
    coords = np.array(t["coords"]).reshape(-1, 3)
    n = coords.shape[0]
    edges = np.array(t["edges"], dtype=np.int32).reshape(-1, 2)
    import naluwind_amd as nw
    c2, conn = nw.hex_box(5, 4, 6, 1.1, 0.15, 3)
    assert np.array_equal(c2, coords)
    e2, ee = nw.hex_edges(conn, n)
    assert np.array_equal(e2, edges)
    areav = nw.hex_edge_area_vector(conn, coords, ee, edges)
    vel = np.array(t["velocity"]).reshape(-1, 3)
    rho = np.array(t["density"])
    top = np.array(t["top"], dtype=np.int32)
    p = np.array(t["pressure"])
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(orc.ContinuityOpts(0.5, 1.0, 1.0, 1.0, 0.0), n, edges,
                            coords, vel, np.array(t["dpdx"]).reshape(-1, 3), rho, p,
                            np.full(n, 2.0), areav, csr=A)
    A.apply_dirichlet(top, 0, 1, p, np.zeros(n))
    S = orc.CsrSystem(n, edges, 1)
    orc.assemble_scalar(orc.ScalarOpts(0.3, 1.0, 1.0, 1.0, 0, orc.Peclet(0, 1.0, 2.0, 4.0)),
                        n, edges, coords, vel, np.array(t["scalar_q"]), np.zeros((n, 3)),
                        rho, np.array(t["diff_flux_coeff"]), areav, np.array(t["mdot"]),
                        csr=S)
    return A, S


def test_two_systems_one_realm_and_solve_leg(out):
    """Two equation systems on one Realm take turns on its context
    (EquationSystem::assemble_and_solve, src/EquationSystem.C:301-350, with
    a GMRES(50) + Jacobi configuration for the continuity system and a
    DirichletBC algorithm): each system equals the oracle's, the solve leg
    ran (iterations, residual recomputed from the downloaded system, the
    reported linear/nonlinear residuals, timers), a call on the inactive
    system throws, and the hypre backend gives the same system and
    solution with its residual reported as relative x ||b||."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "nalu-wind_amd"))
    from tests import problems as pb
    t = out["two_systems"]
    A, S = _oracle_two_systems(t)
    got = _got(t["continuity"])
    pb.assert_csr_close(got, A)
    pb.assert_csr_close(_got(t["scalar"]), S)
    st = t["continuity_solve"]
    assert st["converged"] == 1 and st["iters"] > 0
    assert st["linearSolveIterations"] == st["iters"]
    x = np.array(t["continuity_solution"])
    rp, cols, vals, rhs = got
    Ax = np.array([vals[rp[r]:rp[r + 1]] @ x[cols[rp[r]:rp[r + 1]]]
                   for r in range(len(rp) - 1)])
    r = np.linalg.norm(rhs - Ax)
    assert r <= 1.001e-10 * np.linalg.norm(rhs)
    assert st["linearResidual"] == st["final_abs"]
    assert st["nonLinearResidual"] == pytest.approx(np.linalg.norm(rhs), rel=1e-12)
    assert st["final_abs"] == pytest.approx(r, rel=1e-6, abs=1e-14 * np.linalg.norm(rhs))
    assert all(v > 0 for v in t["timers"]) and st["timerSolve"] > 0
    assert t["inactive_throws"] == 1
    h = _got(t["hypre_continuity"])
    np.testing.assert_array_equal(h[0], got[0])
    np.testing.assert_array_equal(h[1], got[1])
    np.testing.assert_allclose(h[2], got[2], rtol=0, atol=1e-14 * np.abs(got[2]).max())
    np.testing.assert_allclose(h[3], got[3], rtol=0, atol=1e-14 * np.abs(got[3]).max())
    np.testing.assert_allclose(t["hypre_solution"], x, rtol=0,
                               atol=1e-9 * np.abs(x).max())
    hs = t["hypre_solve"]
    assert hs["linearResidual"] == pytest.approx(hs["final_rel"] * hs["bnorm"], rel=1e-15)


def _boundary_sides(nw, coords, types, conn, axis, value):
    """(elem, side) of the faces owned by one element whose nodes all lie
    on the plane coords[:, axis] == value"""
    nsides = {0: 6, 1: 4, 2: 5, 3: 5}
    owner = {}
    for e in range(types.size):
        for s in range(nsides[types[e]]):
            key = tuple(sorted(conn[e, nw.side_nodes(types[e], s)]))
            owner.setdefault(key, []).append((e, s))
    sel = [v[0] for k, v in owner.items() if len(v) == 1 and
           np.all(np.abs(coords[list(k), axis] - value) < 1e-12)]
    return (np.array([e for e, _ in sel], np.int32), np.array([s for _, s in sel], np.int32))


def test_realm_from_exodus_file(tmp_path):
    """Realm::create_mesh on an Exodus II file of a conforming mixed
    Hex8/Tet4/Wed6/Pyr5 box (node set "top", side set "bottom"): the
    device geometry and the continuity system equal the oracle's on the
    mesh read back, the boundary-condition nodes are the plane's nodes, the
    AMG-preconditioned solve reaches its tolerance, and the results file
    holds the realm's fields."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "nalu-wind_amd"))
    import naluwind_amd as nw
    from oracle import orc
    from tests import problems as pb
    coords, types, conn = nw.mixed_box(10, 0.6, 0.3, jitter=0.15, seed=4)
    top = np.where(coords[:, 2] >= 1 - 1e-12)[0].astype(np.int32)
    bottom = _boundary_sides(nw, coords, types, conn, 2, 0.0)
    mesh, res = str(tmp_path / "box.exo"), str(tmp_path / "results.exo")
    order = nw.exo_write(mesh, coords, types, conn, node_sets={"top": top},
                         side_sets={"bottom": bottom})
    r = subprocess.run([EXE, mesh, res], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    t = json.loads(r.stdout)["exodus_realm"]
    types, conn = types[order], conn[order]  # the file's element order
    n = coords.shape[0]
    edges = nw.mixed_edges(types, conn, n)
    assert (t["nnode"], t["nelem"], t["nedge"]) == (n, types.size, edges.shape[0])
    ra = orc.edge_area_vec_mixed(types, conn, coords, edges)
    rv = orc.dual_volume_mixed(types, conn, coords)
    a = np.array(t["area_vector"]).reshape(-1, 3)
    np.testing.assert_allclose(a, ra, rtol=0, atol=1e-14 * np.abs(ra).max())
    np.testing.assert_allclose(t["dual_volume"], rv, rtol=1e-13, atol=0)
    np.testing.assert_array_equal(t["top"], top)
    np.testing.assert_array_equal(t["bottom"], np.where(np.abs(coords[:, 2]) < 1e-12)[0])
    assert t["missing_throws"] == 1
    vel = np.array(t["velocity"]).reshape(-1, 3)
    p = np.array(t["pressure"])
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(orc.ContinuityOpts(0.5, 1.0, 1.0, 1.0, 0.0), n, edges, coords,
                            vel, np.array(t["dpdx"]).reshape(-1, 3), np.array(t["density"]),
                            p, np.full(n, 2.0), ra, csr=A)
    A.apply_dirichlet(top, 0, 1, p, np.zeros(n))
    got = _got(t["continuity"])
    pb.assert_csr_close(got, A, rtol=1e-11)
    st = t["solve"]
    assert st["converged"] == 1 and st["iters"] > 0
    x = np.array(t["solution"])
    res_norm = np.linalg.norm(A.rhs - A.spmv(x))
    assert res_norm <= 1.001e-10 * np.linalg.norm(A.rhs) + 1e-12 * np.abs(A.vals).max()
    with nw.ExodusMesh(res) as out:
        c2, t2, k2, _ = out.read_mesh()
        np.testing.assert_array_equal(c2, coords)
        o2 = np.argsort(types, kind="stable")
        np.testing.assert_array_equal(t2, types[o2])
        np.testing.assert_array_equal(k2, conn[o2])
        assert out.nodal_vars == ["pressure", "velocity_x", "velocity_y", "velocity_z",
                                  "dual_nodal_volume"]
        v, tm = out.nodal_var("velocity_y")
        assert tm == 2.5
        np.testing.assert_array_equal(v, vel[:, 1])
        np.testing.assert_array_equal(out.nodal_var("pressure")[0], p)
        np.testing.assert_array_equal(out.nodal_var("dual_nodal_volume")[0], t["dual_volume"])

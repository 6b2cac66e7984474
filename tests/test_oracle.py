"""CPU checks that pin the oracle: the reference's own golden values, then
closed forms and invariants for what the golds leave open (non-orthogonal
correction, cross-component momentum couplings, geometry closure, GMRES).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import orc
from tests import fixtures as fx

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_golds.json")))

CONT_OPTS = orc.ContinuityOpts(1.0, 1.0, 1.0, 1.0, 0.0)   # dt, gamma1, noc, interp, incompressible
MOM_DEFAULT = orc.MomentumOpts(0.0, 1.0, 1.0, 1.0, 0.0, 0)
SCAL_DEFAULT = orc.ScalarOpts(0.0, 1.0, 1.0, 1.0, 0, orc.Peclet(0, 0.0, 2.0, 4.0))


# ------------------------------------------------------------ golden values

def test_continuity_golds():
    """UnitTestContinuityAdvEdge.C:109-141 (dense mock applier)."""
    f = fx.continuity_fixture()
    g = GOLD["continuity_adv_edge"]
    L, R = orc.assemble_continuity(CONT_OPTS, 8, f["edges"], f["coords"],
                                   f["vel"], f["dpdx"], f["rho"], f["p"],
                                   f["udiag"], f["areav"])
    np.testing.assert_allclose(R, g["rhs"], rtol=0, atol=g["rhs_tol"])
    np.testing.assert_allclose(L.ravel(), g["lhs"], rtol=0, atol=g["lhs_tol"])


def test_momentum_golds():
    """UnitTestMomentumAdvDiffEdge.C:232-266 (mock keeps same-component
    couplings only; 12 sumInto calls = one per edge)."""
    f = fx.momentum_fixture()
    g = GOLD["momentum_adv_diff_edge"]
    assert f["edges"].shape[0] == g["num_sum_into_calls"]
    L, R = orc.assemble_momentum(MOM_DEFAULT, 8, f["edges"], f["coords"],
                                 f["vel"], f["dudx"], f["visc"], f["rho"],
                                 f["areav"], f["mdot"], f["pecfac"],
                                 f["wallmask"])
    np.testing.assert_allclose(R, g["rhs"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(L.ravel(), g["lhs"], rtol=0, atol=g["tol"])


def _scalar_csr():
    f = fx.mixture_fraction_fixture()
    A = orc.CsrSystem(8, f["edges"], 1)
    orc.assemble_scalar(SCAL_DEFAULT, 8, f["edges"], f["coords"], f["vel"],
                        f["q"], f["dqdx"], f["rho"], f["visc"], f["areav"],
                        f["mdot"], csr=A)
    return f, A


def test_scalar_tpetra_golds():
    """UnitTestScalarAdvDiffEdge.C:159-227, serial CSR golds."""
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    _, A = _scalar_csr()
    assert list(A.row_ptr) == g["rowOffsets_serial"]
    assert list(A.cols) == g["cols_serial"]
    np.testing.assert_allclose(A.vals, g["vals_serial"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(A.rhs, g["rhs_serial"], rtol=0, atol=g["tol"])


def test_scalar_fix_pressure_golds():
    """UnitTestScalarAdvDiffEdge.C:229-293: node id 1, p_ref 1, p = density."""
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    f, A = _scalar_csr()
    A.fix_node(g["fix_node_id"] - 1, g["fix_ref_pressure"], f["rho"][0])
    np.testing.assert_allclose(A.vals, g["fixed_vals_serial"], rtol=0,
                               atol=g["tol"])
    np.testing.assert_allclose(A.rhs, g["fixed_rhs_serial"], rtol=0,
                               atol=g["tol"])


def test_scalar_dirichlet_golds():
    """UnitTestScalarAdvDiffEdge.C:295-381: solution 2, bc 0 except node 1."""
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    _, A = _scalar_csr()
    sol = np.full(8, 2.0)
    bc = np.zeros(8)
    bc[0] = 1.0
    A.apply_dirichlet(np.arange(8), 0, 1, sol, bc)
    np.testing.assert_allclose(A.vals, g["dirichlet_vals_serial"], rtol=0,
                               atol=g["tol"])
    np.testing.assert_allclose(A.rhs, g["dirichlet_rhs_serial"], rtol=0,
                               atol=g["tol"])


def test_peclet_golds():
    """UnitTestPecletFunction.C:36-110."""
    g = GOLD["peclet_classic"]
    for pe, fac in zip(g["pe"], g["fac"]):
        assert abs(orc.peclet_classic(g["hybrid"], pe) - fac) < g["tol"]
    g = GOLD["peclet_tanh"]
    for x, fac in zip(g["x"], g["fac"]):
        assert abs(orc.peclet_tanh(g["c1"], g["c2"], x) - fac) < g["tol"]


def _box_edges(n, h):
    coords, conn = fx.hex_box(n, n, n)
    coords = coords * (n * h)
    edges, _ = orc.create_edges_hex(conn, coords.shape[0])
    return coords, edges


def test_peclet_number_uniform_box_closed_form():
    """MomentumEdgePecletAlg.C:74-101 on a uniform box of spacing h with
    uniform u, rho, mu: every edge is axis-aligned, dx = h e_d, so
    Pe = |u_d| h / (mu/rho + 1e-16), computed here from the mesh's own
    edge directions, and the factor is the classic / tanh function of it."""
    h, n = 0.25, 3
    coords, edges = _box_edges(n, h)
    nn = coords.shape[0]
    u = np.array([2.0, -0.5, 0.125])
    rho, mu = 1.6, 4.0e-3
    vel = np.tile(u, (nn, 1))
    for pf, fn in ((orc.Peclet(0, 0.7, 2.0, 4.0), lambda pe: (0.7 * pe) ** 2 / (5 + (0.7 * pe) ** 2)),
                   (orc.Peclet(1, 0.0, 2.0, 4.0), lambda pe: 0.5 * (1 + math.tanh((pe - 2.0) / 4.0)))):
        num, fac = orc.momentum_edge_peclet(edges, coords, vel, np.full(nn, rho),
                                            np.full(nn, mu), pf, with_number=True)
        dx = coords[edges[:, 1]] - coords[edges[:, 0]]
        axis = np.argmax(np.abs(dx), axis=1)
        assert np.allclose(np.abs(dx[np.arange(len(axis)), axis]), h, rtol=1e-14)
        sign = np.sign(dx[np.arange(len(axis)), axis])
        pe = np.abs(u[axis] * sign * h) / (mu / rho + 1e-16)
        np.testing.assert_allclose(num, pe, rtol=1e-14, atol=0)
        np.testing.assert_allclose(fac, [fn(p) for p in pe], rtol=1e-14, atol=1e-300)
    # the three axis values are distinct, so a transposed component shows
    assert len({round(abs(v) * h / (mu / rho), 9) for v in u}) == 3


def test_peclet_number_two_nodes_unequal_properties():
    """Two nodes with unequal rho and mu: diffIp = (muL/rhoL + muR/rhoR)/2
    (MomentumEdgePecletAlg.C:85-86), u.dx with the mean of the two
    velocities (:88-92).  A mean of mu over a mean of rho, or one side's
    ratio, gives a different number here."""
    coords = np.array([[0.1, 0.2, 0.3], [0.6, -0.1, 0.45]])
    edges = np.array([[0, 1]], dtype=np.int32)
    vel = np.array([[1.0, 2.0, -1.0], [3.0, -2.0, 0.5]])
    rho = np.array([0.8, 2.5])
    mu = np.array([3e-3, 1e-2])
    dx = coords[1] - coords[0]
    udotx = 0.5 * dx @ (vel[0] + vel[1])
    diff = 0.5 * (mu[0] / rho[0] + mu[1] / rho[1])
    pe = abs(udotx) / (diff + 1e-16)
    assert abs(abs(udotx) / (0.5 * (mu.sum()) / (0.5 * rho.sum())) - pe) > 1e-3 * pe
    for pf, f in ((orc.Peclet(0, 0.008, 2.0, 4.0), (0.008 * pe) ** 2 / (5 + (0.008 * pe) ** 2)),
                  (orc.Peclet(1, 0.0, 250.0, 100.0), 0.5 * (1 + math.tanh((pe - 250.0) / 100.0)))):
        num, fac = orc.momentum_edge_peclet(edges, coords, vel, rho, mu, pf,
                                            with_number=True)
        assert num[0] == pytest.approx(pe, rel=1e-14)
        assert fac[0] == pytest.approx(f, rel=1e-13)
        assert 0.05 < f < 0.95   # the blend is away from its limits here


def test_scalar_peclet_number_through_the_advective_flux():
    """ScalarEdgeSolverAlg.C:109-110,131-133: the scalar kernel's Peclet
    number uses diffIp = (GammaL/rhoL + GammaR/rhoR)/2 and the mean velocity
    along dx.  With dq/dx = 0, alpha = 0, alpha_upw = 1 and mdot > 0 its
    flux is mdot (f qL + (1-f)(qL+qR)/2) plus the diffusion
    -viscIp |A|^2/(A.dx) (qR - qL), so rhs(L) pins f = f(Pe) in closed
    form, and so does lhs(L,L)."""
    coords = np.array([[0.0, 0.0, 0.0], [0.4, 0.1, -0.05]])
    edges = np.array([[0, 1]], dtype=np.int32)
    vel = np.array([[1.5, -0.4, 0.2], [0.7, 0.3, -0.6]])
    rho = np.array([1.3, 0.6])
    gam = np.array([2e-2, 5e-3])
    q = np.array([0.9, 0.2])
    av = np.array([[0.3, 0.05, 0.02]])
    mdot = np.array([0.8])
    dx = coords[1] - coords[0]
    pe = abs(0.5 * dx @ (vel[0] + vel[1])) / (0.5 * (gam[0] / rho[0] + gam[1] / rho[1]) + 1e-16)
    for hf in (0.03, 0.08):
        f = (hf * pe) ** 2 / (5 + (hf * pe) ** 2)
        assert 0.05 < f < 0.95
        o = orc.ScalarOpts(0.0, 1.0, 1.0, 1.0, 0, orc.Peclet(0, hf, 2.0, 4.0))
        L, R = orc.assemble_scalar(o, 2, edges, coords, vel, q, np.zeros((2, 3)),
                                   rho, gam, av, mdot)
        a = av[0]
        lhsfac = -0.5 * gam.sum() * (a @ a) / (a @ dx)
        diffflux = lhsfac * (q[1] - q[0])
        adv = mdot[0] * (f * q[0] + (1 - f) * 0.5 * (q[0] + q[1]))
        assert R[0] == pytest.approx(-diffflux - adv, rel=1e-13)
        assert R[1] == pytest.approx(diffflux + adv, rel=1e-13)
        assert L[0, 0] == pytest.approx(-lhsfac + mdot[0] * f + 0.5 * mdot[0] * (1 - f),
                                        rel=1e-13)


def test_van_leer_limits():
    # smooth data (equal slopes) -> 1; opposite slopes -> 0; symmetric
    assert orc.van_leer(1.0, 1.0) == pytest.approx(1.0)
    assert orc.van_leer(1.0, -1.0) == 0.0
    assert orc.van_leer(2.0, 1.0) == pytest.approx(orc.van_leer(1.0, 2.0))
    assert orc.van_leer(3.0, 1.0) == pytest.approx(2 * 2 * 3 / 16.0)


def test_geometry_golds():
    """UnitTestGeometryAlg.C:20-100: unit hex dual volume 0.125 per node,
    element volume 1, |A|^2 = 0.0625 on all 12 edges."""
    g = GOLD["geometry_unit_hex"]
    coords, conn, edges, elem_edges, areav = fx.unit_hex()
    dv = orc.dual_volume_hex(conn, coords)
    np.testing.assert_allclose(dv, g["dual_volume"], atol=g["tol"])
    assert abs(dv.sum() - g["element_volume"]) < 1e-15
    assert edges.shape[0] == g["num_edges"]
    np.testing.assert_allclose((areav ** 2).sum(1), g["area_mag_sq"],
                               atol=g["tol"])


# ------------------------------------------------------------ invariants

def _skewed_box(n=3, seed=3):
    coords, conn = fx.hex_box(n, n, n, stretch_z=1.1, jitter=0.2, seed=seed)
    edges, elem_edges = orc.create_edges_hex(conn, coords.shape[0])
    areav = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    dv = orc.dual_volume_hex(conn, coords)
    return coords, conn, edges, areav, dv


def test_geometry_closure_and_volume():
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    # element volumes sum to the box volume (planar box faces)
    assert abs(dv.sum() - 1.0) < 1e-13
    # sum of outward area vectors vanishes around interior control volumes
    S = np.zeros((n, 3))
    np.add.at(S, edges[:, 0], areav)
    np.add.at(S, edges[:, 1], -areav)
    boundary = np.zeros(n, bool)
    for d in range(3):
        lo, hi = coords[:, d].min(), coords[:, d].max()
        boundary |= (np.abs(coords[:, d] - lo) < 1e-12) | (
            np.abs(coords[:, d] - hi) < 1e-12)
    assert np.abs(S[~boundary]).max() < 1e-14
    # uniform box: |A| = h^2 on interior edges, dual volume h^3 inside
    c2, k2 = fx.hex_box(4, 4, 4)
    e2, ee2 = orc.create_edges_hex(k2, c2.shape[0])
    a2 = orc.edge_area_vec_hex(k2, c2, ee2, e2)
    d2 = orc.dual_volume_hex(k2, c2)
    h = 0.25
    assert np.isclose(np.abs(a2).sum(1).max(), h * h)
    assert np.isclose(d2.max(), h ** 3)


def test_nodal_grad_exact_for_linear_interior():
    # exact where each sub-control-surface centroid lies at its edge
    # midpoint in the field's gradient direction: the uniform box
    coords, conn = fx.hex_box(4, 4, 4)
    edges, elem_edges = orc.create_edges_hex(conn, coords.shape[0])
    areav = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    dv = orc.dual_volume_hex(conn, coords)
    g = np.array([0.3, -1.2, 2.5])
    phi = coords @ g + 0.7
    G = orc.nodal_grad_edge(edges, areav, dv, phi)
    boundary = np.zeros(coords.shape[0], bool)
    for d in range(3):
        boundary |= (np.abs(coords[:, d] - coords[:, d].min()) < 1e-12) | (
            np.abs(coords[:, d] - coords[:, d].max()) < 1e-12)
    np.testing.assert_allclose(G[~boundary], np.tile(g, ((~boundary).sum(), 1)),
                               rtol=0, atol=1e-12)


def test_continuity_noc_exact_for_linear_pressure():
    """With noc = 1, a linear pressure with its exact gradient and zero
    velocity gives tmdot = 0 on any (skewed) mesh; noc = 0 does not."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    g = np.array([1.0, -2.0, 0.5])
    p = coords @ g
    args = (coords, np.zeros((n, 3)), np.tile(g, (n, 1)), np.ones(n), p,
            np.full(n, 2.0), areav)
    L, R = orc.assemble_continuity(CONT_OPTS, n, edges, *args)
    assert np.abs(R).max() < 1e-13
    L0, R0 = orc.assemble_continuity(orc.ContinuityOpts(1.0, 1.0, 0.0, 1.0, 0.0),
                                     n, edges, *args)
    assert np.abs(R0).max() > 1e-4
    # LHS is a symmetric graph Laplacian with zero row sums
    assert np.abs(L - L.T).max() < 1e-15
    assert np.abs(L.sum(1)).max() < 1e-13


def test_scalar_diffusion_exact_for_linear_field():
    """Diffusive flux with the non-orthogonal correction equals -mu A.g for a
    linear field with exact gradient; pure-diffusion rows sum to zero."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    g = np.array([0.4, 1.0, -0.3])
    q = coords @ g
    mu = 0.02
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_scalar(SCAL_DEFAULT, n, edges, coords, np.zeros((n, 3)), q,
                        np.tile(g, (n, 1)), np.ones(n), np.full(n, mu), areav,
                        np.zeros(edges.shape[0]), csr=A)
    expect = np.zeros(n)
    flux = -mu * (areav @ g)          # diffFlux per edge
    np.add.at(expect, edges[:, 0], -flux)
    np.add.at(expect, edges[:, 1], flux)
    np.testing.assert_allclose(A.rhs, expect, rtol=0, atol=1e-14)
    D = A.dense()
    assert np.abs(D.sum(1)).max() < 1e-15


def _momentum_rhs(o, n, edges, coords, vel, dudx, visc, areav, mdot, pf, m):
    _, R = orc.assemble_momentum(o, n, edges, coords, vel, dudx, visc,
                                 np.ones(n), areav, mdot, pf, m)
    return R


def test_momentum_lhs_is_jacobian_of_rhs():
    """For frozen gradients/mdot, central advection and no limiter, the
    momentum LHS is -d(RHS)/du (Picard = Newton for this linear flux); this
    pins the cross-component -mu A_i A_j/(A.dx) couplings and the
    relaxation placement that the reference mock does not capture."""
    coords, conn, edges, areav, dv = _skewed_box(2)
    n = coords.shape[0]
    rng = np.random.default_rng(5)
    vel = rng.normal(size=(n, 3))
    dudx = rng.normal(size=(n, 9))
    visc = rng.uniform(0.01, 0.1, n)
    mdot = rng.normal(size=edges.shape[0]) * 0.1
    for pecfac, alpha, aupw in [(0.0, 0.0, 1.0), (1.0, 0.0, 1.0), (0.5, 0.5, 0.3)]:
        o = orc.MomentumOpts(alpha, aupw, 0.0, 1.0, 0.0, 0)
        pf = np.full(edges.shape[0], pecfac)
        mask = np.ones(n)
        A = orc.CsrSystem(n, edges, 3)
        orc.assemble_momentum(o, n, edges, coords, vel, dudx, visc, np.ones(n),
                              areav, mdot, pf, mask, csr=A)
        D = A.dense()
        R0 = A.rhs.copy()
        J = np.zeros((3 * n, 3 * n))
        h = 1e-6
        for k in range(3 * n):
            v2 = vel.copy().ravel()
            v2[k] += h
            A2 = orc.CsrSystem(n, edges, 3)
            orc.assemble_momentum(o, n, edges, coords, v2.reshape(n, 3), dudx,
                                  visc, np.ones(n), areav, mdot, pf, mask,
                                  csr=A2)
            J[:, k] = -(A2.rhs - R0) / h
        np.testing.assert_allclose(D, J, rtol=0, atol=2e-8 * np.abs(D).max())


def test_scalar_lhs_is_jacobian_of_rhs():
    coords, conn, edges, areav, dv = _skewed_box(2)
    n = coords.shape[0]
    rng = np.random.default_rng(7)
    q = rng.normal(size=n)
    dqdx = rng.normal(size=(n, 3))
    mdot = rng.normal(size=edges.shape[0]) * 0.1
    visc = rng.uniform(0.01, 0.1, n)
    for hf in [0.0, 1.0]:
        o = orc.ScalarOpts(0.3, 1.0, 0.0, 1.0, 0, orc.Peclet(0, hf, 2.0, 4.0))
        vel = rng.normal(size=(n, 3)) * 3

        def run(qq):
            A = orc.CsrSystem(n, edges, 1)
            orc.assemble_scalar(o, n, edges, coords, vel, qq, dqdx, np.ones(n),
                                visc, areav, mdot, csr=A)
            return A
        A = run(q)
        D = A.dense()
        J = np.zeros((n, n))
        h = 1e-6
        for k in range(n):
            q2 = q.copy()
            q2[k] += h
            A2 = run(q2)
            J[:, k] = -(A2.rhs - A.rhs) / h
        np.testing.assert_allclose(D, J, rtol=0, atol=2e-8 * np.abs(D).max())


def test_orientation_invariance():
    """Swapping edge node order (with the area vector sign) leaves the CSR
    unchanged when no mdot is exactly zero."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    rng = np.random.default_rng(11)
    vel = rng.normal(size=(n, 3))
    dudx = rng.normal(size=(n, 9))
    visc = rng.uniform(0.01, 0.1, n)
    mdot = orc.mdot_edge(edges, areav, np.ones(n), vel)
    pf = rng.uniform(0, 1, edges.shape[0])
    o = orc.MomentumOpts(0.2, 0.7, 1.0, 0.8, 1.0, 1)
    A = orc.CsrSystem(n, edges, 3)
    orc.assemble_momentum(o, n, edges, coords, vel, dudx, visc, np.ones(n),
                          areav, mdot, pf, np.ones(n), csr=A)
    flip = rng.uniform(size=edges.shape[0]) < 0.5
    e2 = edges.copy()
    e2[flip] = e2[flip][:, ::-1]
    a2 = areav.copy()
    a2[flip] *= -1
    m2 = mdot.copy()
    m2[flip] *= -1
    B = orc.CsrSystem(n, e2, 3)
    orc.assemble_momentum(o, n, e2, coords, vel, dudx, visc, np.ones(n), a2,
                          m2, pf, np.ones(n), csr=B)
    np.testing.assert_allclose(B.vals, A.vals, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(B.rhs, A.rhs, rtol=1e-12, atol=1e-14)


def test_dense_and_csr_agree_same_component():
    """The dense mock and the CSR scatter agree on the couplings the mock
    keeps."""
    f = fx.momentum_fixture()
    L, R = orc.assemble_momentum(MOM_DEFAULT, 8, f["edges"], f["coords"],
                                 f["vel"], f["dudx"], f["visc"], f["rho"],
                                 f["areav"], f["mdot"], f["pecfac"],
                                 f["wallmask"])
    A = orc.CsrSystem(8, f["edges"], 3)
    orc.assemble_momentum(MOM_DEFAULT, 8, f["edges"], f["coords"], f["vel"],
                          f["dudx"], f["visc"], f["rho"], f["areav"], f["mdot"],
                          f["pecfac"], f["wallmask"], csr=A)
    D = A.dense()
    same = np.zeros_like(D, bool)
    for d in range(3):
        same[d::3, d::3] = True
    np.testing.assert_allclose(D[same], L[same], rtol=0, atol=1e-15)
    np.testing.assert_allclose(A.rhs, R, rtol=0, atol=1e-15)


def test_gmres_matches_dense_solve():
    coords, conn, edges, areav, dv = _skewed_box(3)
    n = coords.shape[0]
    rng = np.random.default_rng(2)
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(CONT_OPTS, n, edges, coords, rng.normal(size=(n, 3)),
                            np.zeros((n, 3)), np.ones(n), np.zeros(n),
                            np.ones(n), areav, csr=A)
    top = np.where(np.abs(coords[:, 2] - 1.0) < 1e-12)[0].astype(np.int32)
    A.apply_dirichlet(top, 0, 1, np.zeros(n), np.zeros(n))
    b = A.rhs.copy()
    xd = orc.dense_solve(A.dense(), b)
    for prec in (orc.PREC_NONE, orc.PREC_JACOBI, orc.PREC_CHEBYSHEV):
        x, st = A.gmres(b, kspace=10, max_iters=2000, tol=1e-10, precond=prec)
        assert st.converged
        assert st.final_abs <= 1e-10 * st.bnorm * 1.0001
        np.testing.assert_allclose(x, xd, rtol=0,
                                   atol=1e-7 * np.abs(xd).max())


def _mpnn_system(seed=3):
    coords, conn, edges, areav, dv = _skewed_box(3)
    n = coords.shape[0]
    rng = np.random.default_rng(seed)
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(CONT_OPTS, n, edges, coords, rng.normal(size=(n, 3)),
                            np.zeros((n, 3)), np.ones(n), np.zeros(n),
                            np.ones(n), areav, csr=A)
    top = np.where(np.abs(coords[:, 2] - 1.0) < 1e-12)[0].astype(np.int32)
    A.apply_dirichlet(top, 0, 1, np.zeros(n), np.zeros(n))
    return A, top, rng


def test_mpnn_gradient_matches_finite_differences():
    """The analytic backward of the learned-guess network agrees with
    central differences on every parameter block."""
    A, top, rng = _mpnn_system()
    H, L = 6, 2
    npar = orc.mpnn_nparams(H, L)
    assert npar == H * 5 + H + L * (2 * H * H + H) + 2 * H + 2
    P = rng.normal(size=npar) * 0.4
    w = orc.mpnn_slot_weights(A)
    feat, _ = orc.mpnn_features(A, A.rhs, *rng.normal(size=(5, A.c.nrows)))
    t = rng.normal(size=A.c.nrows)
    loss, g = orc.mpnn_loss_grad(P, H, L, A, w, feat, t)
    d = orc.mpnn_forward(P, H, L, A, w, feat)
    assert loss == pytest.approx(np.mean((d - t) ** 2), rel=1e-14)
    eps = 1e-6
    # every parameter of a small network
    fd = np.empty(npar)
    for i in range(npar):
        p1 = P.copy()
        p1[i] += eps
        p2 = P.copy()
        p2[i] -= eps
        fd[i] = (orc.mpnn_loss_grad(p1, H, L, A, w, feat, t)[0] -
                 orc.mpnn_loss_grad(p2, H, L, A, w, feat, t)[0]) / (2 * eps)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_mpnn_permutation_equivariance():
    """Relabelling the nodes permutes the prediction (the aggregation uses
    the right neighbour indices)."""
    A, top, rng = _mpnn_system(4)
    n = A.c.nrows
    H, L = 5, 3
    P = rng.normal(size=orc.mpnn_nparams(H, L)) * 0.5
    D = A.dense()
    hist = rng.normal(size=(5, n))
    w = orc.mpnn_slot_weights(A)
    feat, _ = orc.mpnn_features(A, A.rhs, *hist)
    d = orc.mpnn_forward(P, H, L, A, w, feat)
    perm = rng.permutation(n)
    Dp = D[np.ix_(perm, perm)]
    B = orc.CsrSystem.from_dense(Dp, A.rhs[perm])
    wp = orc.mpnn_slot_weights(B)
    featp, _ = orc.mpnn_features(B, A.rhs[perm], *hist[:, perm])
    dp = orc.mpnn_forward(P, H, L, B, wp, featp)
    np.testing.assert_allclose(dp, d[perm], rtol=1e-12, atol=1e-12)


def test_mpnn_slot_weights_and_features():
    """Edge weights are symmetric, zero on the diagonal and on Dirichlet
    rows; features follow their closed forms."""
    A, top, rng = _mpnn_system(5)
    n = A.c.nrows
    D = A.dense()
    w = orc.mpnn_slot_weights(A)
    rp, cols = A.row_ptr, A.cols
    W = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            W[r, cols[k]] = w[k]
    np.testing.assert_allclose(W, W.T, rtol=1e-15, atol=0)
    dg = np.abs(np.diag(D))
    ref = np.sqrt(np.abs(D * D.T)) / np.sqrt(np.outer(dg, dg))
    np.fill_diagonal(ref, 0.0)
    np.testing.assert_allclose(W, ref, rtol=1e-14, atol=1e-300)
    assert np.all(W[top] == 0) and np.all(W[:, top] == 0)
    x0 = rng.normal(size=n)
    b = A.rhs.copy()
    h = [x0]
    for _ in range(4):
        h.append(h[-1] + 0.1 * rng.normal(size=n))
    feat, sc = orc.mpnn_features(A, b, *h)
    xb = 3 * h[0] - 3 * h[1] + h[2]
    r = b - D @ xb
    rms = lambda v: np.sqrt(np.mean(v * v))
    sd = rms(h[0] - h[1])
    np.testing.assert_allclose(sc, [rms(b), rms(x0), rms(r), rms(r / dg), sd],
                               rtol=1e-13)
    np.testing.assert_allclose(feat[:, 3], (h[0] - 3 * h[1] + 3 * h[2] - h[3]) / sd,
                               rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(feat[:, 4], (h[0] - 4 * h[1] + 6 * h[2] - 4 * h[3] +
                                            h[4]) / sd, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(orc.mpnn_base_guess(*h), xb, rtol=1e-15, atol=1e-15)
    # shorter histories: the higher differences vanish and the base guess
    # drops to the linear extrapolation / x0; without history the output
    # scale is s_e
    f3h, _ = orc.mpnn_features(A, b, *h[:3])
    assert np.all(f3h[:, 3:] == 0)
    f4h, _ = orc.mpnn_features(A, b, *h[:4])
    assert np.all(f4h[:, 4] == 0) and np.any(f4h[:, 3] != 0)
    np.testing.assert_allclose(orc.mpnn_base_guess(*h[:2]), 2 * h[0] - h[1],
                               rtol=1e-15, atol=1e-15)
    f0, s0 = orc.mpnn_features(A, b, x0)
    assert np.all(f0[:, 3:] == 0) and s0[4] == s0[3]
    assert np.array_equal(orc.mpnn_base_guess(x0), x0)
    np.testing.assert_allclose(feat[:, 0], r / rms(r), rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(feat[:, 1], np.log(dg) - np.log(dg).mean(),
                               rtol=1e-12, atol=1e-14)
    flag = np.zeros(n)
    flag[top] = 1
    np.testing.assert_array_equal(feat[:, 2], flag)


def _bf16(x):
    """Nearest-even bfloat16 of x via FP32 (numpy bit arithmetic)."""
    u = np.array([x], dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return float(u.astype(np.uint32).view(np.float32)[0])


def test_bf16_rounding_helper_closed_forms():
    """The rounding used below: ties go to the even mantissa."""
    assert _bf16(1.0) == 1.0
    assert _bf16(1 + 2 ** -8) == 1.0                      # tie, even below
    assert _bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6          # tie, even above
    assert _bf16(1 + 2 ** -7 + 2 ** -9) == 1 + 2 ** -7    # quarter ulp down
    assert _bf16(2.5078125) == 2.5                        # tie, even below


def test_mpnn_bf16_forward_rounds_where_the_device_rounds():
    """Closed form on a two-node graph whose values sit on bfloat16 ties:
    h0 = bf16(1 + 2^-8) = 1, m = bf16(0.5 h0) = 0.5, W_self = 1 + 2^-7 + 2^-9
    rounds to 1 + 2^-7, W_nbr = 3, and the last layer is not rounded, so
    delta = (1 + 2^-7) + 3 * 0.5 = 2.5078125 exactly.  Rounding the last
    layer would give 2.5, keeping W_self 2.5078125 + 2^-9, keeping h0 or m
    other values again; the FP64 forward gives its own exact sum."""
    D = np.array([[2.0, -1.0], [-1.0, 2.0]])
    A = orc.CsrSystem.from_dense(D, np.zeros(2))
    w = orc.mpnn_slot_weights(A)
    H, L = 16, 1
    P = np.zeros(orc.mpnn_nparams(H, L))
    win, bin_, ws = 0, H * 5, H * 5 + H
    wn, wout = ws + H * H, ws + 2 * H * H + H
    P[win + 0 * 5 + 0] = 1 + 2 ** -8        # W_in[0][0]
    P[ws + 0 * H + 0] = 1 + 2 ** -7 + 2 ** -9  # W_self[0][0]
    P[wn + 0 * H + 0] = 3.0                 # W_nbr[0][0]
    P[wout + 0] = 1.0                       # w_out,1[0]
    feat = np.zeros((2, 5))
    feat[:, 0] = 1.0
    feat[:, 3] = 1.0
    np.testing.assert_array_equal(w[w != 0], 0.5)
    d = orc.mpnn_forward_bf16(P, H, L, A, w, feat)
    np.testing.assert_array_equal(d, [2.5078125, 2.5078125])
    h0 = 1 + 2 ** -8
    exact = (1 + 2 ** -7 + 2 ** -9) * h0 + 3.0 * 0.5 * h0
    np.testing.assert_allclose(orc.mpnn_forward(P, H, L, A, w, feat),
                               [exact, exact], rtol=1e-15)


def test_mpnn_bf16_forward_stays_near_fp64():
    """On a random network the bfloat16 forward differs from the FP64 one
    by the size bfloat16 rounding explains (not more, not zero)."""
    A, top, rng = _mpnn_system(6)
    n = A.c.nrows
    H, L = 32, 3
    P = rng.normal(size=orc.mpnn_nparams(H, L)) * 0.3
    w = orc.mpnn_slot_weights(A)
    feat, _ = orc.mpnn_features(A, A.rhs, *rng.normal(size=(5, n)))
    d = orc.mpnn_forward(P, H, L, A, w, feat)
    db = orc.mpnn_forward_bf16(P, H, L, A, w, feat)
    rel = np.linalg.norm(db - d) / np.linalg.norm(d)
    assert 1e-4 < rel < 3e-2


def test_mdot_alg_is_the_continuity_flux():
    """MdotEdgeAlg and ContinuityEdgeSolverAlg evaluate the same edge flux:
    with dt/gamma1 = tau_s and compressible density scaling (denScale = 1)
    the continuity rhs of node n is sum over its edges of -+mdot_e / tau_s
    (- where n is the edge's left node), for noc on and off and both
    interpolations of rho u."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    rng = np.random.default_rng(21)
    vel, dpdx = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    rho, p = rng.uniform(0.5, 2.0, n), rng.normal(size=n)
    udiag = rng.uniform(1.0, 3.0, n)
    for noc, interp in ((1.0, 1.0), (0.0, 1.0), (1.0, 0.0)):
        dt = 0.3
        o = orc.ContinuityOpts(dt, 1.0, noc, interp, 0.0)
        L, R = orc.assemble_continuity(o, n, edges, coords, vel, dpdx, rho, p,
                                       udiag, areav)
        md = orc.mdot_edge_alg(edges, coords, vel, dpdx, rho, p, udiag, areav,
                               noc, interp)
        ref = np.zeros(n)
        np.add.at(ref, edges[:, 0], -md / dt)
        np.add.at(ref, edges[:, 1], md / dt)
        np.testing.assert_allclose(R, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_mdot_alg_reduces_to_the_unit_test_form():
    """With zero pressure and pressure gradient the projection terms vanish
    and mdot = (rho_L u_L + rho_R u_R)/2 . A, the unit-test form pinned by
    the reference's mdot gold."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    rng = np.random.default_rng(22)
    vel, rho = rng.normal(size=(n, 3)), rng.uniform(0.5, 2.0, n)
    md = orc.mdot_edge_alg(edges, coords, vel, np.zeros((n, 3)), rho,
                           np.zeros(n), np.ones(n), areav, 1.0, 1.0)
    np.testing.assert_allclose(md, orc.mdot_edge(edges, areav, rho, vel),
                               rtol=1e-14, atol=1e-15)


def test_mpnn_residual_loss_gradient_matches_finite_differences():
    """The residual loss mean((D^-1 A u)^2), u = delta - t off the
    Dirichlet rows: its analytic gradient agrees with central differences,
    and it equals the closed form computed with the dense matrix."""
    A, top, rng = _mpnn_system(7)
    H, L = 5, 2
    n = A.c.nrows
    P = rng.normal(size=orc.mpnn_nparams(H, L)) * 0.4
    w = orc.mpnn_slot_weights(A)
    feat, _ = orc.mpnn_features(A, A.rhs, *rng.normal(size=(5, n)))
    t = rng.normal(size=n)
    loss, g = orc.mpnn_loss_grad_residual(P, H, L, A, w, feat, t)
    d = orc.mpnn_forward(P, H, L, A, w, feat)
    D = A.dense()
    u = np.where(feat[:, 2] != 0, 0.0, d - t)
    v = (D @ u) / np.diag(D)
    assert loss == pytest.approx(np.mean(v * v), rel=1e-13)
    assert np.all(v[top] == 0)  # Dirichlet rows carry no residual
    eps = 1e-6
    fd = np.empty(len(P))
    for i in range(len(P)):
        p1, p2 = P.copy(), P.copy()
        p1[i] += eps
        p2[i] -= eps
        fd[i] = (orc.mpnn_loss_grad_residual(p1, H, L, A, w, feat, t)[0] -
                 orc.mpnn_loss_grad_residual(p2, H, L, A, w, feat, t)[0]) / (2 * eps)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)
    # the solution loss is unchanged by the refactoring of the backward
    l0, g0 = orc.mpnn_loss_grad(P, H, L, A, w, feat, t)
    assert l0 == pytest.approx(np.mean((d - t) ** 2), rel=1e-14)


def test_mdot_alg_reference_gold():
    """UnitTestMdotAlg.C:21-78: one unit-cube Hex8, density 1, velocity 10,
    zero pressure and pressure gradient: every edge carries mdot = 2.5."""
    g = GOLD["mdot_edge_unit_hex"]
    coords, conn = fx.hex_box(1, 1, 1)
    n = coords.shape[0]
    edges, elem_edges = orc.create_edges_hex(conn, n)
    areav = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    assert edges.shape[0] == g["num_edges"]
    md = orc.mdot_edge_alg(edges, coords, np.full((n, 3), g["velocity"]),
                           np.full((n, 3), g["dpdx"]), np.full(n, g["density"]),
                           np.full(n, g["pressure"]), np.ones(n), areav, 1.0,
                           g["interp_together"])
    np.testing.assert_allclose(md, g["mdot"], rtol=0, atol=g["tol"])


def test_hypre_skipped_rows_semantics():
    """The hypre backend's skipped rows (HypreLinearSystem.C:2190-2195,
    2326, 1392, 2446-2455): with the check on, the edge assembly leaves the
    rows of skipped nodes at zero and every other row equal to the Tpetra
    assembly; applyDirichletBCs writes diagonal 1 and rhs = bc - sol there;
    a row reset lifts the check, zeroSystem turns it back on."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    rng = np.random.default_rng(5)
    vel = rng.normal(size=(n, 3))
    args = (n, edges, coords, vel, rng.normal(size=(n, 3)), rng.uniform(0.5, 2, n),
            rng.normal(size=n), rng.uniform(0.5, 2, n), areav)
    skipped = np.array([0, 3, n - 1], np.int32)
    full = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(CONT_OPTS, *args, csr=full)
    H = orc.CsrSystem(n, edges, 1)
    H.set_skipped_rows(skipped)
    orc.assemble_continuity(CONT_OPTS, *args, csr=H)
    rows = np.repeat(np.arange(n), np.diff(H.row_ptr))
    sk = np.isin(rows, skipped)
    assert np.all(H.vals[sk] == 0) and np.all(H.rhs[skipped] == 0)
    np.testing.assert_array_equal(H.vals[~sk], full.vals[~sk])
    keep = np.ones(n, bool)
    keep[skipped] = False
    np.testing.assert_array_equal(H.rhs[keep], full.rhs[keep])
    sol, bc = rng.normal(size=n), rng.normal(size=n)
    H.apply_dirichlet_hypre(skipped, sol, bc)
    diag = sk & (H.cols == rows)
    assert np.all(H.vals[diag] == 1.0) and np.all(H.vals[sk & ~diag] == 0)
    np.testing.assert_array_equal(H.rhs[skipped], bc[skipped] - sol[skipped])
    # a block touching a skipped node: its row is left alone, the other row
    # takes its share
    e = edges[np.where((edges[:, 0] == 3) | (edges[:, 1] == 3))[0][0]]
    before = H.vals.copy(), H.rhs.copy()
    H.sum_into(e, np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([5.0, 6.0]))
    i3 = 0 if e[0] == 3 else 1
    assert H.rhs[3] == before[1][3] and H.rhs[e[1 - i3]] == before[1][e[1 - i3]] + [5.0, 6.0][1 - i3]
    # after a reset the check is off: the skipped row takes contributions
    H.reset_rows(np.array([10], np.int32), 0, 1, 1.0, 0.0)
    r3 = H.rhs[3]
    H.sum_into(e, np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([5.0, 6.0]))
    assert H.rhs[3] == r3 + [5.0, 6.0][i3]
    # zeroSystem: check on again
    H.zero()
    orc.assemble_continuity(CONT_OPTS, *args, csr=H)
    assert np.all(H.vals[sk] == 0)


def _hydrostatic_state(n, coords, rng, rho0=1.7, g=(0.3, -0.2, -9.81)):
    """Fluid at rest in hydrostatic balance: p = rho g.x + const, u = 0,
    and dp/dx and the buoyancy source both equal to rho g."""
    g = np.asarray(g)
    p = rho0 * coords @ g + 0.4
    rho = np.full(n, rho0)
    src = np.tile(rho0 * g, (n, 1))
    return g, p, rho, src


def test_balanced_buoyancy_hydrostatic_closed_form():
    """ContinuityEdgeSolverAlg.C:147-152,168-173 and MdotEdgeAlg.C:147-169:
    the balanced buoyancy force is built so that a fluid at rest in
    hydrostatic balance has no mass flux.  On a rectilinear stretched box
    (A parallel to dx on every edge, so the non-orthogonal correction
    vanishes) with p = rho g.x, u = 0 and dp/dx = source = rho g:
      pressure term  -tau (p_R - p_L)|A|^2/(A.dx) = -tau rho g.A,
      gravity term   +tau rho g.A (m_L + m_R)/2,
      G_ip . A       = rho g.A ((1 - m_R)/Ud_R + (1 - m_L)/Ud_L)/2,
    which sum to zero (i) for any masks when Udiag is uniform and (ii) for
    any Udiag when the masks are all 0 or all 1.  A flipped gravity sign, a
    one-sided mask weight, a swapped L/R source or a missing 1/Udiag leaves
    a non-zero flux on some edge.  The continuity rhs and MdotEdgeAlg's
    mdot are both checked."""
    coords, conn = fx.hex_box(4, 3, 5, stretch_z=1.15)
    edges, elem_edges = orc.create_edges_hex(conn, coords.shape[0])
    areav = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    n = coords.shape[0]
    rng = np.random.default_rng(41)
    g, p, rho, src = _hydrostatic_state(n, coords, rng)
    vel = np.zeros((n, 3))
    scale = np.abs(areav).max() * np.abs(rho[0] * g).max() * 2.0
    cases = [(np.full(n, 2.5), rng.uniform(0.0, 1.0, n)),     # (i)
             (rng.uniform(1.0, 4.0, n), np.ones(n)),           # (ii)
             (rng.uniform(1.0, 4.0, n), np.zeros(n))]
    for udiag, mask in cases:
        b = orc.buoyancy(g, src, mask)
        for noc in (1.0, 0.0):
            o = orc.with_buoyancy(orc.ContinuityOpts(0.7, 1.0, noc, 1.0, 0.0), b)
            _, R = orc.assemble_continuity(o, n, edges, coords, vel, src, rho, p,
                                           udiag, areav)
            assert np.abs(R).max() <= 1e-13 * scale / (0.7 * udiag.min())
            md = orc.mdot_edge_alg(edges, coords, vel, src, rho, p, udiag, areav,
                                   noc, 1.0, buoy=b)
            assert np.abs(md).max() <= 1e-13 * scale / udiag.min()
    # not vacuous: the same state without the balanced force (and with the
    # gradient left at zero) carries the hydrostatic flux -tau rho g.A
    udiag = np.full(n, 2.5)
    md0 = orc.mdot_edge_alg(edges, coords, vel, np.zeros((n, 3)), rho, p, udiag,
                            areav, 1.0, 1.0)
    np.testing.assert_allclose(md0, -(1 / 2.5) * rho[0] * areav @ g,
                               rtol=1e-12, atol=1e-14 * scale)
    # and a half-masked state with non-uniform Udiag is not balanced: the
    # closed form of the residual flux is
    # rho g.A [tau (m_L+m_R)/2 - (m_R/Ud_R + m_L/Ud_L)/2]
    udiag = rng.uniform(1.0, 4.0, n)
    mask = rng.uniform(0.0, 1.0, n)
    b = orc.buoyancy(g, src, mask)
    md = orc.mdot_edge_alg(edges, coords, vel, src, rho, p, udiag, areav, 1.0,
                           1.0, buoy=b)
    L, R = edges[:, 0], edges[:, 1]
    tau = 0.5 * (1 / udiag[L] + 1 / udiag[R])
    want = rho[0] * (areav @ g) * (tau * 0.5 * (mask[L] + mask[R]) -
                                   0.5 * (mask[R] / udiag[R] + mask[L] / udiag[L]))
    np.testing.assert_allclose(md, want, rtol=1e-11, atol=1e-13 * scale)
    assert np.abs(want).max() > 1e-3 * scale


def test_balanced_buoyancy_masked_off_is_the_plain_flux():
    """With buoyancy_source_mask = 0 the balanced force adds nothing: the
    continuity system and mdot equal those without it, on a jittered mesh
    with random fields (the source is not read through any other path)."""
    coords, conn, edges, areav, dv = _skewed_box()
    n = coords.shape[0]
    rng = np.random.default_rng(42)
    vel, dpdx = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    rho, p = rng.uniform(0.5, 2.0, n), rng.normal(size=n)
    udiag = rng.uniform(1.0, 3.0, n)
    b = orc.buoyancy((0.1, 0.2, -9.8), rng.normal(size=(n, 3)), np.zeros(n))
    o = orc.ContinuityOpts(0.3, 1.0, 1.0, 1.0, 0.0)
    L0, R0 = orc.assemble_continuity(o, n, edges, coords, vel, dpdx, rho, p,
                                     udiag, areav)
    L1, R1 = orc.assemble_continuity(orc.with_buoyancy(o, b), n, edges, coords,
                                     vel, dpdx, rho, p, udiag, areav)
    assert np.array_equal(L0, L1) and np.array_equal(R0, R1)
    m0 = orc.mdot_edge_alg(edges, coords, vel, dpdx, rho, p, udiag, areav)
    m1 = orc.mdot_edge_alg(edges, coords, vel, dpdx, rho, p, udiag, areav, buoy=b)
    assert np.array_equal(m0, m1)


def test_mpnn_history_differences_extrapolate_polynomials_exactly():
    """The base guess and the two history features pin the stencils by the
    polynomials they extrapolate exactly: with a history h_j = p(-j) of a
    polynomial p in time, the base guess is p(1) for quadratics, and the
    heads a = (1, 0) and a = (1, 1), which add s_d f_3 and s_d (f_3 + f_4),
    give p(1) for cubics and quartics (a wrong stencil coefficient, a
    swapped history index or a missing s_d breaks one of these)."""
    A, top, rng = _mpnn_system(8)
    n = A.c.nrows
    c = rng.normal(size=(5, n))
    t = np.arange(0, -5, -1, dtype=float)
    for deg in (2, 3, 4):
        hist = [sum(c[k] * tj ** k for k in range(deg + 1)) for tj in t]
        p1 = sum(c[k] for k in range(deg + 1))   # p(1)
        feat, sc = orc.mpnn_features(A, A.rhs, *hist)
        xb = orc.mpnn_base_guess(*hist)
        if deg == 2:
            np.testing.assert_allclose(xb, p1, rtol=1e-12, atol=1e-11)
            assert np.abs(feat[:, 3:]).max() < 1e-10
        pred = xb + sc[4] * (feat[:, 3] + (feat[:, 4] if deg == 4 else 0.0))
        np.testing.assert_allclose(pred, p1, rtol=1e-11, atol=1e-10)
        if deg >= 3:   # the base guess alone is not exact
            assert np.abs(xb - p1).max() > 1e-3

"""GPU parity of the edge solver algorithms through the C ABI: the reference
unit-test golds, then the oracle on seeded skewed meshes over the scheme
options, determinism, the sumInto / resetRows / Dirichlet / fix-node row
operations and error paths."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import orc
from tests import fixtures as fx
from tests import problems as pb

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_golds.json")))


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


@pytest.fixture
def ctx(nw):
    c = nw.Context(0)
    yield c
    c.close()


def _dense(got, n):
    rp, cols, vals, rhs = got
    D = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            D[r, cols[k]] += vals[k]
    return D, rhs


def _load_fixture(nw, ctx, f, extra):
    ctx.mesh_set(f["coords"].shape[0], f["edges"])
    ctx.field_upload(nw.F_COORDINATES, f["coords"])
    ctx.field_upload(nw.E_AREA_VECTOR, f["areav"])
    for fid, key in extra:
        ctx.field_upload(fid, f[key])


# ------------------------------------------------------------ golds

def test_continuity_golds(nw, ctx):
    f = fx.continuity_fixture()
    _load_fixture(nw, ctx, f, [(nw.F_VELOCITY, "vel"), (nw.F_DPDX, "dpdx"),
                               (nw.F_DENSITY, "rho"), (nw.F_PRESSURE, "p"),
                               (nw.F_UDIAG, "udiag")])
    ctx.build_graph(1)
    ctx.assemble_continuity(nw.continuity_defaults())
    D, rhs = _dense(ctx.download(), 8)
    g = GOLD["continuity_adv_edge"]
    np.testing.assert_allclose(rhs, g["rhs"], rtol=0, atol=g["rhs_tol"])
    np.testing.assert_allclose(D.ravel(), g["lhs"], rtol=0, atol=g["lhs_tol"])


def test_momentum_golds(nw, ctx):
    f = fx.momentum_fixture()
    _load_fixture(nw, ctx, f, [(nw.F_VELOCITY, "vel"), (nw.F_DUDX, "dudx"),
                               (nw.F_VISCOSITY, "visc"), (nw.F_DENSITY, "rho"),
                               (nw.E_MASS_FLOW_RATE, "mdot"),
                               (nw.E_PECLET_FACTOR, "pecfac"),
                               (nw.F_WALL_MASK, "wallmask")])
    ctx.build_graph(3)
    ctx.assemble_momentum(nw.momentum_defaults())
    D, rhs = _dense(ctx.download(), 24)
    g = GOLD["momentum_adv_diff_edge"]
    np.testing.assert_allclose(rhs, g["rhs"], rtol=0, atol=g["tol"])
    # the reference mock keeps same-component couplings only
    G = np.array(g["lhs"]).reshape(24, 24)
    same = np.zeros((24, 24), bool)
    for d in range(3):
        same[d::3, d::3] = True
    np.testing.assert_allclose(D[same], G[same], rtol=0, atol=g["tol"])


def _scalar_fixture(nw, ctx):
    f = fx.mixture_fraction_fixture()
    _load_fixture(nw, ctx, f, [(nw.F_VELOCITY, "vel"), (nw.F_SCALAR_Q, "q"),
                               (nw.F_DQDX, "dqdx"), (nw.F_DENSITY, "rho"),
                               (nw.F_DIFF_FLUX_COEFF, "visc"),
                               (nw.E_MASS_FLOW_RATE, "mdot")])
    ctx.build_graph(1)
    ctx.assemble_scalar(nw.scalar_defaults())
    return f


def test_scalar_tpetra_golds(nw, ctx):
    _scalar_fixture(nw, ctx)
    rp, cols, vals, rhs = ctx.download()
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    assert list(rp) == g["rowOffsets_serial"]
    assert list(cols) == g["cols_serial"]
    np.testing.assert_allclose(vals, g["vals_serial"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(rhs, g["rhs_serial"], rtol=0, atol=g["tol"])


def test_scalar_fix_pressure_golds(nw, ctx):
    f = _scalar_fixture(nw, ctx)
    ctx.field_upload(nw.F_PRESSURE, f["rho"])  # the test fixes "density"
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    ctx.fix_pressure_at_node(g["fix_node_id"] - 1, g["fix_ref_pressure"])
    rp, cols, vals, rhs = ctx.download()
    np.testing.assert_allclose(vals, g["fixed_vals_serial"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(rhs, g["fixed_rhs_serial"], rtol=0, atol=g["tol"])


def test_scalar_dirichlet_golds(nw, ctx):
    _scalar_fixture(nw, ctx)
    ctx.field_upload(nw.F_SCALAR_Q, np.full(8, 2.0))
    bc = np.zeros(8)
    bc[0] = 1.0
    # any 1-wide node field can hold the bc values (the reference test uses
    # its viscosity field the same way)
    ctx.field_upload(nw.F_DIFF_FLUX_COEFF, bc)
    ctx.apply_dirichlet(np.arange(8), nw.F_SCALAR_Q, nw.F_DIFF_FLUX_COEFF, 0, 1)
    rp, cols, vals, rhs = ctx.download()
    g = GOLD["scalar_adv_diff_edge_tpetra"]
    np.testing.assert_allclose(vals, g["dirichlet_vals_serial"], rtol=0, atol=g["tol"])
    np.testing.assert_allclose(rhs, g["dirichlet_rhs_serial"], rtol=0, atol=g["tol"])


# ------------------------------------------------------------ oracle parity

@pytest.fixture(scope="module")
def skewed():
    return pb.Problem(9, 7, 8, stretch=1.15, jitter=0.25, seed=4)


CONT_CASES = [dict(), dict(noc=0.0), dict(interp_together=0.0),
              dict(incompressible=1.0, dt=0.3, gamma1=1.5)]


@pytest.mark.parametrize("kw", CONT_CASES)
def test_continuity_matches_oracle(nw, ctx, skewed, kw):
    o = nw.continuity_defaults(**kw)
    skewed.upload(ctx)
    ctx.build_graph(1)
    ctx.assemble_continuity(o)
    pb.assert_csr_close(ctx.download(), skewed.oracle_continuity(pb.to_orc_continuity(o)))


def _upload_buoyancy(nw, ctx, P, seed, mask=None):
    rng = np.random.default_rng(seed)
    src = rng.normal(size=(P.n, 3))
    if mask is None:
        mask = rng.uniform(0.0, 1.0, P.n)
    ctx.field_upload(nw.F_BUOYANCY_SOURCE, src)
    ctx.field_upload(nw.F_BUOYANCY_SOURCE_MASK, mask)
    return src, mask


@pytest.mark.parametrize("kw", [dict(), dict(noc=0.0, interp_together=0.0),
                                dict(incompressible=1.0, dt=0.3, gamma1=1.5)])
def test_continuity_balanced_buoyancy_matches_oracle(nw, ctx, skewed, kw):
    """ContinuityEdgeSolverAlg with use_balanced_buoyancy_force_
    (ContinuityEdgeSolverAlg.C:147-152,168-173) on the jittered box, random
    source and mask in [0, 1], against the oracle."""
    g = (0.3, -0.7, -9.81)
    o = nw.continuity_defaults(**kw)
    o.balanced_buoyancy = 1
    o.gravity = (ctypes.c_double * 3)(*g)
    skewed.upload(ctx)
    src, mask = _upload_buoyancy(nw, ctx, skewed, 17)
    ctx.build_graph(1)
    ctx.assemble_continuity(o)
    b = orc.buoyancy(g, src, mask)
    A = orc.CsrSystem(skewed.n, skewed.edges, 1)
    orc.assemble_continuity(orc.with_buoyancy(pb.to_orc_continuity(o), b),
                            skewed.n, skewed.edges, skewed.coords, skewed.vel,
                            skewed.dpdx, skewed.rho, skewed.p, skewed.udiag,
                            skewed.areav, csr=A)
    pb.assert_csr_close(ctx.download(), A)
    # MdotEdgeAlg with the same force
    ctx.compute_mdot(kw.get("noc", 1.0), kw.get("interp_together", 1.0), gravity=g)
    got = ctx.field_download(nw.E_MASS_FLOW_RATE)
    want = orc.mdot_edge_alg(skewed.edges, skewed.coords, skewed.vel, skewed.dpdx,
                             skewed.rho, skewed.p, skewed.udiag, skewed.areav,
                             kw.get("noc", 1.0), kw.get("interp_together", 1.0),
                             buoy=b)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_balanced_buoyancy_hydrostatic_rest_on_device(nw, ctx):
    """A fluid at rest in hydrostatic balance on a rectilinear stretched box
    (p = rho g.x, u = 0, dp/dx = source = rho g, any mask, uniform Udiag):
    the device continuity rhs and mdot vanish to rounding, as the oracle's
    closed form fixes, while the same state without the force carries the
    flux -tau rho g.A."""
    P = pb.Problem(10, 8, 9, stretch=1.12, seed=5)
    g = np.array([0.3, -0.2, -9.81])
    rho0 = 1.3
    P.vel[:] = 0.0
    P.rho[:] = rho0
    P.p = rho0 * P.coords @ g + 0.25
    P.dpdx = np.tile(rho0 * g, (P.n, 1))
    P.udiag[:] = 2.0
    P.upload(ctx)
    ctx.field_upload(nw.F_BUOYANCY_SOURCE, P.dpdx)
    ctx.field_upload(nw.F_BUOYANCY_SOURCE_MASK,
                     np.random.default_rng(3).uniform(0.0, 1.0, P.n))
    ctx.build_graph(1)
    o = nw.continuity_defaults(dt=0.5)
    o.balanced_buoyancy = 1
    o.gravity = (ctypes.c_double * 3)(*g)
    ctx.assemble_continuity(o)
    rhs = ctx.download()[3]
    scale = np.abs(P.areav).max() * rho0 * np.abs(g).max()
    assert np.abs(rhs).max() <= 1e-12 * scale
    ctx.compute_mdot(gravity=g)
    assert np.abs(ctx.field_download(nw.E_MASS_FLOW_RATE)).max() <= 1e-12 * scale
    ctx.field_upload(nw.F_DPDX, np.zeros((P.n, 3)))
    ctx.compute_mdot()
    np.testing.assert_allclose(ctx.field_download(nw.E_MASS_FLOW_RATE),
                               -0.5 * rho0 * P.areav @ g, rtol=1e-12,
                               atol=1e-13 * scale)


SCAL_CASES = [dict(), dict(alpha=0.5, alpha_upw=0.3, upw=0.7, relax=0.8),
              dict(use_limiter=1, alpha=1.0),
              dict(peclet=("classic", 1.0)), dict(peclet=("tanh", 0.0))]


def _scal_opts(nw, kw):
    kw = dict(kw)
    o = nw.scalar_defaults()
    if "peclet" in kw:
        form, hf = kw.pop("peclet")
        o.peclet = nw.Peclet(0, hf, 2.0, 4.0) if form == "classic" else nw.Peclet(1, 0.0, 2.0, 4.0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("kw", SCAL_CASES)
def test_scalar_matches_oracle(nw, ctx, skewed, kw):
    o = _scal_opts(nw, kw)
    skewed.upload(ctx)
    ctx.build_graph(1)
    ctx.assemble_scalar(o)
    pb.assert_csr_close(ctx.download(), skewed.oracle_scalar(pb.to_orc_scalar(o)))


MOM_CASES = [dict(), dict(use_limiter=1, include_divU=1.0),
             dict(alpha=0.3, alpha_upw=0.6, upw=0.5, relax=0.7)]


@pytest.mark.parametrize("kw", MOM_CASES)
def test_momentum_matches_oracle(nw, ctx, skewed, kw):
    o = nw.momentum_defaults(**kw)
    skewed.upload(ctx)
    ctx.build_graph(3)
    ctx.assemble_momentum(o)
    pb.assert_csr_close(ctx.download(), skewed.oracle_momentum(pb.to_orc_momentum(o)))


@pytest.fixture(scope="module")
def mixed():
    """Conforming mixed Hex8/Tet4/Wed6/Pyr5 box (60/30/10 % of the cells cut
    Hex8/Tet4/Wed6, transition cells between), jittered, nodes permuted then
    RCM-numbered: irregular rows."""
    return pb.Problem(9, 0, 0, jitter=0.15, seed=6, mixed=(0.6, 0.3))


@pytest.mark.parametrize("kw", CONT_CASES[:2])
def test_mixed_mesh_continuity_matches_oracle(nw, ctx, mixed, kw):
    o = nw.continuity_defaults(**kw)
    mixed.upload(ctx)
    ctx.build_graph(1)
    ctx.assemble_continuity(o)
    pb.assert_csr_close(ctx.download(), mixed.oracle_continuity(pb.to_orc_continuity(o)))


def test_mixed_mesh_scalar_and_momentum_match_oracle(nw, ctx, mixed):
    mixed.upload(ctx)
    ctx.build_graph(1)
    o = _scal_opts(nw, dict(alpha=0.5, use_limiter=1))
    ctx.assemble_scalar(o)
    pb.assert_csr_close(ctx.download(), mixed.oracle_scalar(pb.to_orc_scalar(o)))
    ctx.build_graph(3)
    m = nw.momentum_defaults(alpha=0.5)
    ctx.assemble_momentum(m)
    pb.assert_csr_close(ctx.download(), mixed.oracle_momentum(pb.to_orc_momentum(m)))


def test_limiter_of_nearly_flat_slopes_matches_oracle(nw, ctx, skewed):
    """The van Leer limiter (EdgeKernelUtils.h:18-24) of slopes of opposite
    sign is exactly 0; with the gradients scaled to 1e-9 the projected
    slopes nearly cancel the jump, so the limiter's denominator is near its
    eps and any residue in the numerator (the rounding error of the product,
    left by a fused multiply-add) would become a limiter of order one."""
    import copy
    P = copy.copy(skewed)
    P.dqdx = skewed.dqdx * 1e-9
    P.dudx = skewed.dudx * 1e-9
    P.upload(ctx)
    ctx.build_graph(1)
    o = _scal_opts(nw, dict(alpha=0.5, use_limiter=1))
    ctx.assemble_scalar(o)
    pb.assert_csr_close(ctx.download(), P.oracle_scalar(pb.to_orc_scalar(o)))
    ctx.build_graph(3)
    m = nw.momentum_defaults(alpha=0.5, use_limiter=1)
    ctx.assemble_momentum(m)
    pb.assert_csr_close(ctx.download(), P.oracle_momentum(pb.to_orc_momentum(m)))


def test_uniform_32_cubed_scalar_config(nw, ctx):
    """32^3 uniform box, Taylor-Green velocity, mu = 1e-3, seeded phi with
    its edge nodal gradient: the scalar workload, against the oracle."""
    P = pb.Problem(32, 32, 32, seed=7)
    x, y, z = P.coords.T
    P.vel = np.stack([np.sin(x) * np.cos(y) * np.cos(z),
                      -np.cos(x) * np.sin(y) * np.cos(z), 0 * z], 1)
    P.rho = np.ones(P.n)
    P.dflux = np.full(P.n, 1e-3)
    P.q = np.random.default_rng(7).uniform(0, 1, P.n)
    P.dqdx = pb.nw().nodal_grad_edge(P.edges, P.areav, P.dualvol, P.q)
    P.mdot = pb.nw().mdot_edge(P.edges, P.areav, P.rho, P.vel)
    P.upload(ctx)
    ctx.build_graph(1)
    o = nw.scalar_defaults()
    ctx.assemble_scalar(o)
    pb.assert_csr_close(ctx.download(), P.oracle_scalar(pb.to_orc_scalar(o)))


def test_rows_spanning_chunks(nw, ctx):
    """Rows longer than a 64-lane chunk: a star graph whose hub has 150
    neighbours, so the hub row spans three chunks."""
    rng = np.random.default_rng(3)
    n = 151
    edges = np.stack([np.zeros(150, np.int32), np.arange(1, 151, dtype=np.int32)], 1)
    coords = rng.normal(size=(n, 3))
    coords[0] = 0
    areav = coords[1:] * rng.uniform(0.5, 1.5, (150, 1))
    fields = dict(vel=rng.normal(size=(n, 3)), dpdx=rng.normal(size=(n, 3)),
                  rho=rng.uniform(0.5, 2, n), p=rng.normal(size=n),
                  udiag=rng.uniform(0.5, 2, n))
    ctx.mesh_set(n, edges)
    ctx.field_upload(nw.F_COORDINATES, coords)
    ctx.field_upload(nw.E_AREA_VECTOR, areav)
    for fid, k in [(nw.F_VELOCITY, "vel"), (nw.F_DPDX, "dpdx"), (nw.F_DENSITY, "rho"),
                   (nw.F_PRESSURE, "p"), (nw.F_UDIAG, "udiag")]:
        ctx.field_upload(fid, fields[k])
    ctx.build_graph(1)
    o = nw.continuity_defaults()
    ctx.assemble_continuity(o)
    A = orc.CsrSystem(n, edges, 1)
    orc.assemble_continuity(pb.to_orc_continuity(o), n, edges, coords, fields["vel"],
                            fields["dpdx"], fields["rho"], fields["p"],
                            fields["udiag"], areav, csr=A)
    pb.assert_csr_close(ctx.download(), A)


def test_deterministic_and_accumulate(nw, ctx, skewed):
    skewed.upload(ctx)
    ctx.build_graph(3)
    o = nw.momentum_defaults(use_limiter=1)
    ctx.assemble_momentum(o)
    a = ctx.download()
    ctx.assemble_momentum(o)
    b = ctx.download()
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    o.accumulate = 1
    ctx.assemble_momentum(o)
    c = ctx.download()
    assert np.array_equal(c[2], 2 * a[2]) and np.array_equal(c[3], 2 * a[3])


def test_momentum_peclet_matches_oracle(nw, ctx, skewed):
    skewed.upload(ctx)
    for pf in [nw.Peclet(0, 1.0, 2.0, 4.0), nw.Peclet(1, 0.0, 5.0, 3.0)]:
        ctx.momentum_edge_peclet(pf)
        got = ctx.field_download(nw.E_PECLET_FACTOR)
        ref = orc.momentum_edge_peclet(skewed.edges, skewed.coords, skewed.vel,
                                       skewed.rho, skewed.visc,
                                       orc.Peclet(pf.form, pf.hybrid, pf.c1, pf.c2))
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("pf_args", [(0, 1.0, 2.0, 4.0), (1, 0.0, 5.0, 3.0)])
@pytest.mark.parametrize("problem", ["skewed", "mixed"])
def test_momentum_with_peclet_equals_the_two_algorithms(nw, ctx, skewed, mixed,
                                                         pf_args, problem):
    """The fused pass (nw_assemble_momentum_with_peclet) gives the system and
    the edge Peclet number and factor of MomentumEdgePecletAlg followed by
    MomentumEdgeSolverAlg, and both match the oracle."""
    P = skewed if problem == "skewed" else mixed
    pf = nw.Peclet(*pf_args)
    o = nw.momentum_defaults(use_limiter=1)
    P.upload(ctx)
    ctx.build_graph(3)
    ctx.momentum_edge_peclet(pf)
    ctx.assemble_momentum(o)
    two = ctx.download()
    pn2 = ctx.field_download(nw.E_PECLET_NUMBER)
    pf2 = ctx.field_download(nw.E_PECLET_FACTOR)
    ctx.field_fill(nw.E_PECLET_NUMBER, -1.0)
    ctx.field_fill(nw.E_PECLET_FACTOR, -1.0)
    ctx.assemble_momentum_with_peclet(o, pf)
    one = ctx.download()
    pn1 = ctx.field_download(nw.E_PECLET_NUMBER)
    pf1 = ctx.field_download(nw.E_PECLET_FACTOR)
    np.testing.assert_allclose(pn1, pn2, rtol=1e-15, atol=0)
    np.testing.assert_allclose(pf1, pf2, rtol=1e-15, atol=1e-300)
    assert np.array_equal(one[0], two[0]) and np.array_equal(one[1], two[1])
    scale = np.abs(two[2]).max()
    np.testing.assert_allclose(one[2], two[2], rtol=0, atol=1e-14 * scale)
    np.testing.assert_allclose(one[3], two[3], rtol=0,
                               atol=1e-14 * np.abs(two[3]).max())
    ref_n, ref = orc.momentum_edge_peclet(P.edges, P.coords, P.vel, P.rho, P.visc,
                                          orc.Peclet(pf.form, pf.hybrid, pf.c1, pf.c2),
                                          with_number=True)
    np.testing.assert_allclose(pf1, ref, rtol=1e-13, atol=1e-15)
    # u.dx can cancel: bound each edge by the rounding of its terms
    # sum |dx_d (uL_d + uR_d) / 2| / diffIp
    L, R = P.edges[:, 0], P.edges[:, 1]
    dx = P.coords[R] - P.coords[L]
    diff_ip = 0.5 * (P.visc[L] / P.rho[L] + P.visc[R] / P.rho[R])
    mag = np.abs(0.5 * dx * (P.vel[L] + P.vel[R])).sum(1) / diff_ip
    for pn in (pn1, pn2):
        assert (np.abs(pn - ref_n) <= 1e-13 * np.abs(ref_n) + 1e-15 * mag).all()


# ------------------------------------------------------------ row operations

def test_sum_into_reset_and_dirichlet_match_oracle(nw, ctx, skewed):
    skewed.upload(ctx)
    ctx.build_graph(3)
    o = nw.momentum_defaults()
    ctx.assemble_momentum(o)
    A = skewed.oracle_momentum(pb.to_orc_momentum(o))
    rng = np.random.default_rng(0)
    e = skewed.edges[5]
    lhs = rng.normal(size=(6, 6))
    rhs = rng.normal(size=6)
    ctx.sum_into(e, lhs, rhs)
    A.sum_into(e, lhs, rhs)
    nodes = np.array([3, 10, 11], np.int32)
    ctx.reset_rows(nodes, 1, 3, 2.5, -1.0)
    A.reset_rows(nodes, 1, 3, 2.5, -1.0)
    sol = rng.normal(size=(skewed.n, 3))
    bc = rng.normal(size=(skewed.n, 3))
    ctx.field_upload(nw.F_VELOCITY, sol)
    ctx.field_upload(nw.F_BC_VALUE, bc)
    dn = np.array([0, 1, 40], np.int32)
    ctx.apply_dirichlet(dn, nw.F_VELOCITY, nw.F_BC_VALUE, 0, 3)
    A.apply_dirichlet(dn, 0, 3, sol, bc)
    pb.assert_csr_close(ctx.download(), A)


def test_sum_into_missing_column_is_an_error(nw, ctx, skewed):
    skewed.upload(ctx)
    ctx.build_graph(1)
    before = ctx.download()
    far = np.array([0, skewed.n - 1], np.int32)   # not an edge
    with pytest.raises(nw.NwError) as ei:
        ctx.sum_into(far, np.ones((2, 2)), np.ones(2))
    assert ei.value.code == nw.E_NOGRAPHCOL
    after = ctx.download()
    assert np.array_equal(before[2], after[2])


def test_bad_calls_are_errors(nw, ctx, skewed):
    with pytest.raises(nw.NwError):
        ctx.assemble_continuity(nw.continuity_defaults())  # no mesh/graph
    skewed.upload(ctx)
    ctx.build_graph(1)
    with pytest.raises(nw.NwError):
        ctx.assemble_momentum(nw.momentum_defaults())  # 1-dof graph
    with pytest.raises(nw.NwError):
        ctx.field_upload(nw.F_VELOCITY, np.zeros(5))
    with pytest.raises(nw.NwError):
        ctx.mesh_set(4, np.array([[0, 0]], np.int32))


def test_nonfinite_input_is_an_error(nw, ctx, skewed):
    """With the finite check on, a NaN in an input field (pressure of one
    node, for continuity; velocity, for momentum) makes the assembly return
    E_NONFINITE naming a row next to that node; finite inputs pass and the
    check changes no value."""
    P = skewed
    P.upload(ctx)
    ctx.build_graph(1)
    ctx.assemble_continuity(nw.continuity_defaults())
    ref = ctx.download()
    ctx.set_check_finite(True)
    ctx.assemble_continuity(nw.continuity_defaults())
    got = ctx.download()
    assert all(np.array_equal(a, b) for a, b in zip(ref, got))
    bad = 57
    nbrs = set(P.edges[(P.edges == bad).any(1)].ravel().tolist())
    p = P.p.copy()
    p[bad] = np.nan
    ctx.field_upload(nw.F_PRESSURE, p)
    with pytest.raises(nw.NwError) as ei:
        ctx.assemble_continuity(nw.continuity_defaults())
    assert ei.value.code == nw.E_NONFINITE
    row = int(str(ei.value).split("node ")[1].split()[0])
    assert row in nbrs
    ctx.field_upload(nw.F_PRESSURE, P.p)
    ctx.build_graph(3)
    v = P.vel.copy()
    v[bad, 1] = np.inf
    ctx.field_upload(nw.F_VELOCITY, v)
    with pytest.raises(nw.NwError) as ei:
        ctx.assemble_momentum(nw.momentum_defaults())
    assert ei.value.code == nw.E_NONFINITE
    ctx.set_check_finite(False)
    ctx.assemble_momentum(nw.momentum_defaults())  # unchecked: no error


def test_schedule_order_does_not_change_results(nw, monkeypatch):
    """The chunk order of the row schedule (forced here by NW_ROW_ORDER)
    only decides which rows run together: every order gives the same
    continuity and momentum systems and nodal gradient, bit for bit, and
    they match the oracle."""
    P = pb.Problem(16, 16, 16, stretch=1.1, jitter=0.2, seed=9)
    oc = nw.continuity_defaults()
    om = nw.momentum_defaults(use_limiter=1)
    out = {}
    for order in ("natural", "morton", "sweep"):
        monkeypatch.setenv("NW_ROW_ORDER", order)
        res = []
        for ndof in (1, 3):
            c = nw.Context(0)
            try:
                P.upload(c)
                c.build_graph(ndof)
                if ndof == 1:
                    c.assemble_continuity(oc)
                    res.append(c.download())
                    c.nodal_grad(nw.F_VELOCITY, nw.F_DUDX)
                    res.append(c.field_download(nw.F_DUDX))
                else:
                    c.assemble_momentum(om)
                    res.append(c.download())
            finally:
                c.close()
        out[order] = res
    ref = out["natural"]
    for order in ("morton", "sweep"):
        for a, b in zip(ref, out[order]):
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y), order
            else:
                assert np.array_equal(a, b), order
    pb.assert_csr_close(ref[0], P.oracle_continuity(pb.to_orc_continuity(oc)))
    pb.assert_csr_close(ref[2], P.oracle_momentum(pb.to_orc_momentum(om)))


def test_edge_order_and_orientation_do_not_change_the_system(nw, skewed):
    """STK hands edges in its own order, node[0] not always the smaller id;
    the area vector and the mass flow rate follow the edge's orientation.
    A permuted edge list with half the edges flipped (their area vectors
    and mass flow rates negated) gives the same three systems."""
    import types
    P = skewed
    rng = np.random.default_rng(2)
    perm = rng.permutation(P.ne)
    flip = rng.random(P.ne) < 0.5
    sgn = np.where(flip, -1.0, 1.0)
    e2 = P.edges[perm].copy()
    e2[flip] = e2[flip][:, ::-1]
    node_fields = [(nw.F_COORDINATES, P.coords), (nw.F_VELOCITY, P.vel),
                   (nw.F_DENSITY, P.rho), (nw.F_PRESSURE, P.p), (nw.F_DPDX, P.dpdx),
                   (nw.F_UDIAG, P.udiag), (nw.F_VISCOSITY, P.visc),
                   (nw.F_DUDX, P.dudx), (nw.F_WALL_MASK, P.wall),
                   (nw.F_SCALAR_Q, P.q), (nw.F_DQDX, P.dqdx),
                   (nw.F_DIFF_FLUX_COEFF, P.dflux)]

    def run(edges, areav, mdot, pecfac, ndof, fn):
        c = nw.Context(0)
        try:
            c.mesh_set(P.n, edges)
            for fid, a in node_fields:
                c.field_upload(fid, a)
            c.field_upload(nw.E_AREA_VECTOR, areav)
            c.field_upload(nw.E_MASS_FLOW_RATE, mdot)
            c.field_upload(nw.E_PECLET_FACTOR, pecfac)
            c.build_graph(ndof)
            fn(c)
            return c.download()
        finally:
            c.close()

    cases = [(1, lambda c: c.assemble_continuity(nw.continuity_defaults())),
             (1, lambda c: c.assemble_scalar(nw.scalar_defaults(use_limiter=1))),
             (3, lambda c: c.assemble_momentum(nw.momentum_defaults(use_limiter=1)))]
    for ndof, fn in cases:
        a = run(P.edges, P.areav, P.mdot, P.pecfac, ndof, fn)
        b = run(e2, P.areav[perm] * sgn[:, None], P.mdot[perm] * sgn,
                P.pecfac[perm], ndof, fn)
        ref = types.SimpleNamespace(row_ptr=a[0], cols=a[1], vals=a[2], rhs=a[3])
        pb.assert_csr_close(b, ref, rtol=1e-13)


# ------------------------------------------------------------ hypre backend

@pytest.mark.parametrize("ndof", [1, 3])
def test_skipped_rows_match_oracle(nw, ctx, skewed, ndof):
    """nw_linsys_set_skipped_rows: the hypre applier's skipped rows
    (HypreLinearSystem.C:2165-2239, 2326, 2408-2456) through the edge
    assembly (overwrite and accumulate), sumInto, applyDirichletBCs, a row
    reset that lifts the check and zeroSystem that restores it, against the
    oracle's restatement."""
    P = skewed
    P.upload(ctx)
    ctx.build_graph(ndof)
    rng = np.random.default_rng(31)
    skipped = np.unique(rng.integers(0, P.n, 25)).astype(np.int32)
    ctx.set_skipped_rows(skipped)
    if ndof == 1:
        o = nw.continuity_defaults(accumulate=1)
        run = lambda: ctx.assemble_continuity(o)
        oracle = lambda A: orc.assemble_continuity(
            pb.to_orc_continuity(o), P.n, P.edges, P.coords, P.vel, P.dpdx, P.rho,
            P.p, P.udiag, P.areav, csr=A)
        sol_f, bc_f = nw.F_PRESSURE, nw.F_PRESSURE_BC
        sol, bc = P.p, rng.normal(size=P.n)
    else:
        o = nw.momentum_defaults(accumulate=1, alpha=0.3)
        run = lambda: ctx.assemble_momentum(o)
        oracle = lambda A: orc.assemble_momentum(
            pb.to_orc_momentum(o), P.n, P.edges, P.coords, P.vel, P.dudx, P.visc,
            P.rho, P.areav, P.mdot, P.pecfac, P.wall, csr=A)
        sol_f, bc_f = nw.F_VELOCITY, nw.F_BC_VALUE
        sol, bc = P.vel, rng.normal(size=(P.n, 3))
    ctx.field_upload(bc_f, bc)
    A = orc.CsrSystem(P.n, P.edges, ndof)
    A.set_skipped_rows(skipped)
    ctx.zero()
    run()
    oracle(A)
    pb.assert_csr_close(ctx.download(), A)
    # overwrite mode writes the skipped rows as zeros
    o.accumulate = 0
    ctx.field_upload(sol_f, sol)
    run()
    pb.assert_csr_close(ctx.download(), A)
    o.accumulate = 1
    ctx.apply_dirichlet(skipped, sol_f, bc_f, 0, ndof)
    A.apply_dirichlet_hypre(skipped, sol, bc)
    pb.assert_csr_close(ctx.download(), A)
    # a second assembly adds to the other rows only
    run()
    oracle(A)
    pb.assert_csr_close(ctx.download(), A)
    e = P.edges[np.where(np.isin(P.edges, skipped).any(1))[0][0]]
    lhs = rng.normal(size=(2 * ndof, 2 * ndof))
    rhs = rng.normal(size=2 * ndof)
    ctx.sum_into(e, lhs, rhs)
    A.sum_into(e, lhs, rhs)
    pb.assert_csr_close(ctx.download(), A)
    # a row reset lifts the check: the skipped rows take contributions
    nodes = np.array([int(np.setdiff1d(np.arange(P.n), skipped)[0])], np.int32)
    ctx.reset_rows(nodes, 0, ndof, 2.0, 0.5)
    A.reset_rows(nodes, 0, ndof, 2.0, 0.5)
    ctx.sum_into(e, lhs, rhs)
    A.sum_into(e, lhs, rhs)
    run()
    oracle(A)
    pb.assert_csr_close(ctx.download(), A)
    # zeroSystem turns it back on
    ctx.zero()
    A.zero()
    run()
    oracle(A)
    pb.assert_csr_close(ctx.download(), A)
    ctx.set_skipped_rows(np.zeros(0, np.int32))   # cleared: the Tpetra system
    ctx.zero()
    run()
    pb.assert_csr_close(ctx.download(), _tpetra(P, ndof, oracle))


def _tpetra(P, ndof, oracle):
    A = orc.CsrSystem(P.n, P.edges, ndof)
    oracle(A)
    return A


def test_momentum_extract_diagonal(nw, ctx, skewed):
    """NGPApplyCoeff::extract_diagonal (src/SolverAlgorithm.C:89-105): the
    (0, 0) entry of each node's (node, node) block of this assembly is added
    to momentum_diag, in overwrite and accumulate modes and in the one-pass
    form; other calls leave the field alone."""
    P = skewed
    P.upload(ctx)
    ctx.build_graph(3)
    u0 = np.random.default_rng(2).uniform(1.0, 2.0, P.n)
    ctx.field_upload(nw.F_UDIAG, u0)
    o = nw.momentum_defaults(alpha=0.2, relax=0.7, extract_diagonal=1)
    A = P.oracle_momentum(pb.to_orc_momentum(o))
    D = A.dense()
    d00 = np.diag(D)[0::3]
    ctx.assemble_momentum(o)
    pb.assert_csr_close(ctx.download(), A)
    u1 = ctx.field_download(nw.F_UDIAG)
    np.testing.assert_allclose(u1, u0 + d00, rtol=1e-13, atol=0)
    o.accumulate = 1
    ctx.assemble_momentum(o)
    np.testing.assert_allclose(ctx.field_download(nw.F_UDIAG), u0 + 2 * d00,
                               rtol=1e-13, atol=0)
    o.extract_diagonal = 0
    o.accumulate = 0
    ctx.field_upload(nw.F_UDIAG, u0)
    ctx.assemble_momentum(o)
    np.testing.assert_array_equal(ctx.field_download(nw.F_UDIAG), u0)
    o.extract_diagonal = 1
    pf = nw.Peclet(0, 1.0, 2.0, 4.0)
    ctx.momentum_edge_peclet(pf)
    ctx.assemble_momentum(o)
    u2 = ctx.field_download(nw.F_UDIAG)
    ctx.field_upload(nw.F_UDIAG, u0)
    ctx.assemble_momentum_with_peclet(o, pf)
    np.testing.assert_allclose(ctx.field_download(nw.F_UDIAG), u2, rtol=1e-13, atol=0)


# ------------------------------------------------------------ device sumInto

@pytest.mark.parametrize("ndof", [1, 3])
def test_device_sum_into_matches_host_and_oracle(nw, skewed, ndof):
    """nw_linsys_sum_into_device (the device-resident CoeffApplier): many
    edge and self blocks, edges repeated so that entries take several
    contributions, from device arrays.  Two contexts hold the same
    assembled system; one takes each round of blocks through the host
    sumInto, the other through the device one: the systems stay bitwise
    equal (same block-order sums) and equal to the oracle's.  Rounds of
    decreasing size reuse the scratch.  A block whose column is not in the
    graph, or whose node id is out of range, is an error that applies
    nothing."""
    P = skewed
    rng = np.random.default_rng(40 + ndof)
    o = nw.continuity_defaults() if ndof == 1 else nw.momentum_defaults(alpha=0.2)
    ctxs = [nw.Context(0), nw.Context(0)]
    try:
        for c in ctxs:
            P.upload(c)
            c.build_graph(ndof)
            if ndof == 1:
                c.assemble_continuity(o)
            else:
                c.assemble_momentum(o)
        A = (P.oracle_continuity(pb.to_orc_continuity(o)) if ndof == 1 else
             P.oracle_momentum(pb.to_orc_momentum(o)))
        for nb, nent in ((3000, 2), (500, 1), (40, 2)):
            if nent == 2:
                nodes = P.edges[rng.integers(0, P.ne, nb)].copy()
                nodes[::2] = nodes[::2, ::-1]   # both orientations
            else:
                nodes = rng.integers(0, P.n, (nb, 1)).astype(np.int32)
            nr = nent * ndof
            lhs = rng.normal(size=(nb, nr, nr))
            rhs = rng.normal(size=(nb, nr))
            ctxs[0].sum_into(nodes, lhs, rhs)
            bufs = [nw.DeviceBuffer(a) for a in (nodes.astype(np.int32), lhs, rhs)]
            ctxs[1].sum_into_device(*bufs, nb, nent)
            host, dev = ctxs[0].download(), ctxs[1].download()
            for k in range(4):
                assert np.array_equal(dev[k], host[k]), k
            for b in range(nb):
                A.sum_into(nodes[b], lhs[b], rhs[b])
            pb.assert_csr_close(dev, A)
            for d in bufs:
                d.free()
        c = ctxs[1]
        lhs = nw.DeviceBuffer(np.ones((2, 2 * ndof, 2 * ndof)))
        rhs = nw.DeviceBuffer(np.ones((2, 2 * ndof)))
        before = c.download()
        far = nw.DeviceBuffer(np.array([[0, 1], [0, P.n - 1]], np.int32))
        with pytest.raises(nw.NwError) as ei:
            c.sum_into_device(far, lhs, rhs, 2, 2)
        assert ei.value.code == nw.E_NOGRAPHCOL
        bad = nw.DeviceBuffer(np.array([[0, P.n]], np.int32))
        with pytest.raises(nw.NwError) as ei:
            c.sum_into_device(bad, lhs, rhs, 1, 2)
        assert ei.value.code == nw.E_ARG
        after = c.download()
        assert all(np.array_equal(x, y) for x, y in zip(before, after))
    finally:
        for c in ctxs:
            c.close()


def test_device_sum_into_skips_hypre_rows(nw, ctx, skewed):
    """With skipped rows set (and their check on) the device sumInto drops
    their contributions, as the host one and the oracle do."""
    P = skewed
    P.upload(ctx)
    ctx.build_graph(1)
    skipped = np.array([0, 5, 17, P.n - 1], np.int32)
    ctx.set_skipped_rows(skipped)
    A = orc.CsrSystem(P.n, P.edges, 1)
    A.set_skipped_rows(skipped)
    rng = np.random.default_rng(9)
    nodes = P.edges[np.isin(P.edges, skipped).any(1)][:50].astype(np.int32)
    nodes = np.concatenate([nodes, P.edges[:100]]).astype(np.int32)
    lhs = rng.normal(size=(len(nodes), 2, 2))
    rhs = rng.normal(size=(len(nodes), 2))
    ctx.zero()
    ctx.sum_into_device(nw.DeviceBuffer(nodes), nw.DeviceBuffer(lhs),
                        nw.DeviceBuffer(rhs), len(nodes), 2)
    for b in range(len(nodes)):
        A.sum_into(nodes[b], lhs[b], rhs[b])
    pb.assert_csr_close(ctx.download(), A)

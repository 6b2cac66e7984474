"""Mixed Hex8/Tet4/Wed6 meshes (the unstructured configuration): product
geometry against the oracle, the discrete Gauss identity of the CVFEM
sub-control surfaces, closure on face-conforming meshes, and the
generator's element mix and numbering."""
import numpy as np
import pytest

from oracle import orc
from tests import problems as pb


@pytest.fixture(scope="module")
def nw():
    return pb.nw()


def _rand_tet(rng):
    c = rng.normal(size=(4, 3))
    if np.dot(np.cross(c[1] - c[0], c[2] - c[0]), c[3] - c[0]) < 0:
        c[[1, 2]] = c[[2, 1]]
    return c


def _gauss(areas, lr, c):
    return sum(np.outer(areas[k], c[r] - c[l]) for k, (l, r) in enumerate(lr))


def test_element_gauss_identity():
    """sum_scs A (x_R - x_L)^T = V I holds exactly for linear tets and for
    prisms with parallel end faces (the CVFEM dual is exact for linear
    fields there); a transposed table, wrong sign or misplaced sub-point
    breaks it."""
    rng = np.random.default_rng(0)
    for _ in range(5):
        c = _rand_tet(rng)
        v = np.dot(np.cross(c[1] - c[0], c[2] - c[0]), c[3] - c[0]) / 6
        M = _gauss(orc.tet_scs_areav(c), orc.TET_LRSCV, c)
        np.testing.assert_allclose(M / v, np.eye(3), atol=1e-12)
        b = rng.normal(size=(3, 3))
        h = rng.normal(size=3)
        if np.dot(np.cross(b[1] - b[0], b[2] - b[0]), h) < 0:
            h = -h
        w = np.vstack([b, b + h])
        v = abs(np.dot(np.cross(b[1] - b[0], b[2] - b[0]), h)) / 2
        M = _gauss(orc.wed_scs_areav(w), orc.WED_LRSCV, w)
        np.testing.assert_allclose(M / v, np.eye(3), atol=1e-12)
    # the pinned Hex8 geometry satisfies it on a parallelepiped too
    E = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1],
                     [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    c = base @ E
    M = _gauss(orc.hex_scs_areav(c), orc.HEX_LRSCV, c)
    np.testing.assert_allclose(M / abs(np.linalg.det(E)), np.eye(3), atol=1e-12)


def test_rigid_motion_and_scaling():
    rng = np.random.default_rng(1)
    c = _rand_tet(rng)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    a = orc.tet_scs_areav(c)
    np.testing.assert_allclose(orc.tet_scs_areav(2.5 * c @ Q.T + 1.0),
                               6.25 * a @ Q.T, atol=1e-12)


_HEX_REF = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
_WED_REF = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)


def _affine(rng, ref):
    A = rng.normal(size=(3, 3))
    if np.linalg.det(A) < 0:
        A[:, 0] = -A[:, 0]
    return ref @ A.T + rng.normal(size=3), np.linalg.det(A)


def test_scv_volumes_split_affine_elements_evenly():
    """Affine maps keep volume ratios, and on the reference elements the
    median-dual cut gives every node the same share (a third of a triangle
    per corner, half a prism's height), so every sub-control volume of an
    affine tetrahedron is V/4, of an affine prism V/6 and of a
    parallelepiped V/8 (their sub-volumes are planar-faced, where Grandy's
    formula is exact).  A wrong or swapped point in a sub-volume table, or a
    misassigned node, breaks the equality."""
    rng = np.random.default_rng(3)
    for _ in range(5):
        c = _rand_tet(rng)
        v = np.dot(np.cross(c[1] - c[0], c[2] - c[0]), c[3] - c[0]) / 6
        np.testing.assert_allclose(orc.tet_scv_volume(c), v / 4, rtol=1e-12)
        p, det = _affine(rng, _WED_REF)
        np.testing.assert_allclose(orc.wed_scv_volume(p), det / 12, rtol=1e-12)
        h, det = _affine(rng, _HEX_REF)
        np.testing.assert_allclose(orc.hex_scv_volume(h), det / 8, rtol=1e-12)


def test_scv_volumes_follow_their_nodes():
    """Relabelling the element's vertices by a symmetry of its topology
    (a rotation of the tetrahedron's vertices, the prism's end swap) permutes
    the sub-volumes the same way, on a distorted element."""
    rng = np.random.default_rng(4)
    c = _rand_tet(rng)
    perm = [1, 2, 0, 3]  # rotation about vertex 3 keeps the orientation
    np.testing.assert_allclose(orc.tet_scv_volume(c[perm]),
                               orc.tet_scv_volume(c)[perm], rtol=1e-12)
    p = _affine(rng, _WED_REF)[0] + 0.05 * rng.normal(size=(6, 3))
    rot = [1, 2, 0, 4, 5, 3]
    np.testing.assert_allclose(orc.wed_scv_volume(p[rot]),
                               orc.wed_scv_volume(p)[rot], rtol=1e-12)


@pytest.mark.parametrize("mix", [(0.6, 0.3), (0.0, 1.0), (0.0, 0.0)])
def test_mixed_dual_volume_fills_the_box(nw, mix):
    """Without jitter every face is planar and conforming, so the dual
    volumes tile the unit box; each is positive."""
    c, t, conn = nw.mixed_box(6, mix[0], mix[1], 0.0, seed=5)
    v = orc.dual_volume_mixed(t, conn, c)
    assert (v > 0).all()
    assert abs(v.sum() - 1.0) < 1e-13


@pytest.mark.parametrize("mix", [(0.6, 0.3), (0.0, 1.0), (0.0, 0.0), (1.0, 0.0)])
def test_mixed_mesh_geometry(nw, mix):
    c, t, conn = nw.mixed_box(7, mix[0], mix[1], 0.0, seed=5)
    e = nw.mixed_edges(t, conn, c.shape[0])
    a = nw.mixed_edge_area_vector(t, conn, c, e)
    np.testing.assert_allclose(a, orc.edge_area_vec_mixed(t, conn, c, e),
                               rtol=0, atol=1e-15)
    # whole-mesh Gauss identity: sum_e A_e dx_e^T = |box| I (no jitter)
    M = np.einsum("ei,ej->ij", a, c[e[:, 1]] - c[e[:, 0]])
    np.testing.assert_allclose(M, np.eye(3), atol=1e-12)
    # closure at interior nodes when faces conform (single topology)
    if mix in [(0.0, 1.0), (0.0, 0.0), (1.0, 0.0)]:
        clo = np.zeros_like(c)
        np.add.at(clo, e[:, 0], a)
        np.add.at(clo, e[:, 1], -a)
        inner = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
        assert np.abs(clo[inner]).max() < 1e-15


def test_mixed_generator(nw):
    n = 10
    c, t, conn = nw.mixed_box(n, 0.6, 0.3, 0.1, seed=2)
    assert c.shape[0] == (n + 1) ** 3
    cells = np.bincount(t, minlength=3) / np.array([1, 6, 2])
    np.testing.assert_array_equal(cells, [600, 300, 100])
    # every node is used and appears once per coordinate lattice point
    used = np.unique(conn[conn >= 0])
    np.testing.assert_array_equal(used, np.arange(c.shape[0]))
    # positive tet volumes
    tets = conn[t == 1][:, :4]
    x = c[tets]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]),
                    x[:, 3] - x[:, 0])
    assert (vol > 0).all()
    # reverse Cuthill-McKee keeps the bandwidth far below a random numbering
    e = nw.mixed_edges(t, conn, c.shape[0])
    assert np.abs(e[:, 1] - e[:, 0]).max() < 0.2 * c.shape[0]
    # deterministic
    c2, t2, conn2 = nw.mixed_box(n, 0.6, 0.3, 0.1, seed=2)
    assert np.array_equal(c, c2) and np.array_equal(conn, conn2)

"""Mixed Hex8/Tet4/Wed6/Pyr5 meshes (the unstructured configuration):
product geometry against the oracle, the discrete Gauss identity of the
CVFEM sub-control surfaces, the Pyr5 sub-control geometry against closed
forms, closure on the conforming generator's meshes, and the generator's
element mix, face matching and numbering."""
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


_PYR_REF = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0],
                     [0, 0, 1]], float)  # UnitTestUtils.C:226-235
_PYR_FACES = [[3, 2, 1, 0], [0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]


def _pyr_volume(c):
    """Pyramid with a planar base: base vector area . (apex - base point) / 3."""
    base = 0.5 * np.cross(c[2] - c[0], c[3] - c[1])
    return np.dot(base, c[4] - c[0]) / 3


def test_pyramid_sub_control_surfaces_close_every_sub_volume():
    """Each node's sub-control volume is bounded by its element-face share
    (the quads corner / edge midpoint / face centroid / edge midpoint of
    the faces at that node, outward) and by the sub-control surfaces of its
    edges (outward from L, inward to R): a closed surface, whose vector area
    vanishes for any vertex positions, planar faces or not.  A wrong point,
    a swapped L/R pair or a reversed surface in the Pyr5 tables breaks it."""
    rng = np.random.default_rng(11)
    for k in range(6):
        c = _PYR_REF + (0 if k == 0 else rng.normal(scale=0.25, size=(5, 3)))
        a = orc.pyr_scs_areav(c)
        res = np.zeros((5, 3))
        for ip, (l, r) in enumerate(orc.PYR_LRSCV):
            res[l] += a[ip]
            res[r] -= a[ip]
        for F in _PYR_FACES:
            cen = c[F].mean(0)
            for j, v in enumerate(F):
                q = [c[v], 0.5 * (c[v] + c[F[(j + 1) % len(F)]]), cen,
                     0.5 * (c[v] + c[F[j - 1]])]
                res[v] += 0.5 * np.cross(q[2] - q[0], q[3] - q[1])
        assert np.abs(res).max() < 1e-14 * (1 + np.abs(a).max())


def test_pyramid_sub_control_volumes():
    """On the reference pyramid (UnitTestUtils.C:226-235, volume 4/3) the
    apex sub-volume is convex, so its volume is that of the convex hull of
    its ten vertices (apex, edge midpoints and triangle centroids of the
    four sides, element centroid): 52/135; the four base sub-volumes are
    equal by symmetry and fill the rest, 32/135 each.  Under affine maps all
    five scale by the determinant, and on any planar-based pyramid they sum
    to its volume."""
    from scipy.spatial import ConvexHull
    c = _PYR_REF
    apex_pts = np.vstack([c[4], c.mean(0)] +
                         [0.5 * (c[i] + c[4]) for i in range(4)] +
                         [(c[i] + c[(i + 1) % 4] + c[4]) / 3 for i in range(4)])
    hull = ConvexHull(apex_pts).volume
    assert abs(hull - 52 / 135) < 1e-14
    v = orc.pyr_scv_volume(c)
    np.testing.assert_allclose(v, [32 / 135] * 4 + [52 / 135], rtol=1e-14)
    rng = np.random.default_rng(12)
    for _ in range(5):
        p, det = _affine(rng, _PYR_REF)
        np.testing.assert_allclose(orc.pyr_scv_volume(p), det * v, rtol=1e-12)
        np.testing.assert_allclose(orc.pyr_scv_volume(p).sum(), _pyr_volume(p),
                                   rtol=1e-13)
    # the Gauss identity holds on affine images (linear fields are exact)
    p, det = _affine(rng, _PYR_REF)
    M = _gauss(orc.pyr_scs_areav(p), orc.PYR_LRSCV, p)
    np.testing.assert_allclose(M / (det * 4 / 3), np.eye(3), atol=1e-12)
    # a rotation of the base about the apex permutes the sub-volumes
    q = p + 0.05 * rng.normal(size=(5, 3))
    rot = [1, 2, 3, 0, 4]
    np.testing.assert_allclose(orc.pyr_scv_volume(q[rot]),
                               orc.pyr_scv_volume(q)[rot], rtol=1e-12)


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
    # closure at interior nodes: every face is shared by matching elements
    clo = np.zeros_like(c)
    np.add.at(clo, e[:, 0], a)
    np.add.at(clo, e[:, 1], -a)
    inner = np.all((c > 1e-9) & (c < 1 - 1e-9), axis=1)
    assert np.abs(clo[inner]).max() < 1e-15


_FACES = {0: [[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3],
              [0, 3, 2, 1], [4, 5, 6, 7]],
          1: [[0, 1, 3], [1, 2, 3], [0, 3, 2], [0, 2, 1]],
          2: [[0, 1, 4, 3], [1, 2, 5, 4], [0, 3, 5, 2], [0, 2, 1], [3, 4, 5]],
          3: [[0, 1, 4], [1, 2, 4], [2, 3, 4], [0, 4, 3], [0, 3, 2, 1]]}


def test_mixed_generator(nw):
    n = 10
    c, t, conn = nw.mixed_box(n, 0.6, 0.3, 0.1, seed=2)
    cnt = np.bincount(t, minlength=4)
    # nodes: the lattice plus one centre per transition cell, each cell
    # accounted for: a transition cell has a pyramid or two tetrahedra on
    # each of its 6 faces, and Kuhn cells (never converted) keep f_tet n^3
    nx = c.shape[0] - (n + 1) ** 3
    cut = 6 * nx - cnt[3]
    tet_cells = (cnt[1] - 2 * cut) / 6
    assert tet_cells == round(0.3 * n ** 3)
    assert cnt[0] + cnt[2] / 2 + tet_cells + nx == n ** 3
    assert 0 < nx < n ** 3 and cnt[3] > 0
    # every node is used
    used = np.unique(conn[conn >= 0])
    np.testing.assert_array_equal(used, np.arange(c.shape[0]))
    # positive tetrahedra; pyramid bases face their apex
    x = c[conn[t == 1][:, :4]]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]),
                    x[:, 3] - x[:, 0])
    assert (vol > 0).all()
    x = c[conn[t == 3][:, :5]]
    base = np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 1])
    assert (np.einsum("ij,ij->i", base, x[:, 4] - x.mean(1)) > 0).all()
    # conforming: every face is shared by exactly two elements with the same
    # node set, except the faces on the box boundary, held by one
    seen = {}
    for el in range(t.size):
        for F in _FACES[int(t[el])]:
            key = tuple(sorted(conn[el, F]))
            seen[key] = seen.get(key, 0) + 1
    assert max(seen.values()) == 2
    lattice = np.round(c * n)
    for key, k in seen.items():
        if k == 1:  # all nodes on one box side
            pts = lattice[list(key)]
            assert any(np.all(pts[:, d] == b) for d in range(3) for b in (0, n))
    # reverse Cuthill-McKee keeps the bandwidth far below a random numbering
    e = nw.mixed_edges(t, conn, c.shape[0])
    assert np.abs(e[:, 1] - e[:, 0]).max() < 0.2 * c.shape[0]
    # deterministic
    c2, t2, conn2 = nw.mixed_box(n, 0.6, 0.3, 0.1, seed=2)
    assert np.array_equal(c, c2) and np.array_equal(conn, conn2)

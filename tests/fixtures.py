"""Test inputs that restate the reference unit-test fixtures, plus small
structured-box meshes for parity runs at oracle-friendly sizes.

Reference fixtures restated here:
  TestKernelHex8Mesh::fill_mesh_and_init_fields
      unit_tests/kernels/UnitTestKernelUtils.h:315-342  (one Hex8, unit cube)
  LowMachKernelHex8Mesh / ContinuityKernelHex8Mesh   :376-475
  MomentumKernelHex8Mesh / MomentumEdgeHex8Mesh       :484-595
  MixtureFractionKernelHex8Mesh                       :1357-1464
  TrigFieldFunction (a = 0.3)      unit_tests/kernels/UnitTestKernelUtils.C:31-367
  mixture_fraction / property functions               :667-714
  calc_mass_flow_rate                                 :725-769
"""
import math

import numpy as np

from oracle import orc

A_TRIG = 0.3


def unit_hex():
    """STK "generated:1x1x1": node ids 1..8 numbered x fastest, one Hex8 in
    exodus node order; edges from the element edge table, node[0] = smaller
    id."""
    coords = np.array([(x, y, z) for z in (0.0, 1.0) for y in (0.0, 1.0)
                       for x in (0.0, 1.0)])
    conn = np.array([[0, 1, 3, 2, 4, 5, 7, 6]], np.int32)
    edges, elem_edges = orc.create_edges_hex(conn, 8)
    areav = orc.edge_area_vec_hex(conn, coords, elem_edges, edges)
    return coords, conn, edges, elem_edges, areav


def trig_velocity(c, a=A_TRIG):
    x, y = c[:, 0], c[:, 1]
    v = np.zeros_like(c)
    v[:, 0] = -np.cos(a * math.pi * x) * np.sin(a * math.pi * y)
    v[:, 1] = np.sin(a * math.pi * x) * np.cos(a * math.pi * y)
    return v


def trig_dudx(c, a=A_TRIG):
    x, y = c[:, 0], c[:, 1]
    ap = a * math.pi
    g = np.zeros((c.shape[0], 9))
    g[:, 0] = ap * np.sin(ap * x) * np.sin(ap * y)
    g[:, 1] = -ap * np.cos(ap * x) * np.cos(ap * y)
    g[:, 3] = ap * np.cos(ap * x) * np.cos(ap * y)
    g[:, 4] = -ap * np.sin(ap * x) * np.sin(ap * y)
    return g


def trig_pressure(c, a=A_TRIG):
    x, y = c[:, 0], c[:, 1]
    return -0.25 * (np.cos(2 * a * math.pi * x) + np.cos(2 * a * math.pi * y))


def trig_dpdx(c, a=A_TRIG):
    x, y = c[:, 0], c[:, 1]
    g = np.zeros_like(c)
    g[:, 0] = 0.5 * a * math.pi * np.sin(2 * a * math.pi * x)
    g[:, 1] = 0.5 * a * math.pi * np.sin(2 * a * math.pi * y)
    return g


def mixture_fraction(c, znot, amf):
    return (znot * np.cos(amf * math.pi * c[:, 0]) *
            np.cos(amf * math.pi * c[:, 1]) * np.cos(amf * math.pi * c[:, 2]))


def mdot_three_component_field(mdot):
    """The unit-test fixtures declare the edge field "mass_flow_rate" with 3
    components per edge but calc_mass_flow_rate stores edge ie's value at
    flat index ie of the bucket, while the solver algorithms read component
    0 of each edge (flat index 3*e).  The reference golds were produced with
    that layout, so the fixture reproduces it."""
    ne = mdot.shape[0]
    flat = np.zeros(3 * ne)
    flat[:ne] = mdot
    return flat[0::3].copy()


def continuity_fixture():
    coords, conn, edges, elem_edges, areav = unit_hex()
    n = coords.shape[0]
    return dict(coords=coords, edges=edges, areav=areav, vel=trig_velocity(coords),
                dpdx=trig_dpdx(coords), rho=np.ones(n), p=trig_pressure(coords),
                udiag=np.ones(n))


def momentum_fixture():
    coords, conn, edges, elem_edges, areav = unit_hex()
    n = coords.shape[0]
    vel = trig_velocity(coords)
    rho = np.ones(n)
    mdot = mdot_three_component_field(orc.mdot_edge(edges, areav, rho, vel))
    return dict(coords=coords, edges=edges, areav=areav, vel=vel,
                dudx=trig_dudx(coords), visc=np.full(n, 0.1), rho=rho,
                mdot=mdot, pecfac=np.zeros(edges.shape[0]),
                wallmask=np.ones(n))


def mixture_fraction_fixture():
    coords, conn, edges, elem_edges, areav = unit_hex()
    n = coords.shape[0]
    # the fixture passes (amf_, znot_) = (2, 1) into (znot, amf)
    z = mixture_fraction(coords, znot=2.0, amf=1.0)
    rho = 1.0 / (z / 0.163 + (1.0 - z) / 1.18)
    visc = 1.967e-5 * z + 1.85e-5 * (1.0 - z)
    vel = trig_velocity(coords)
    mdot = mdot_three_component_field(orc.mdot_edge(edges, areav, rho, vel))
    return dict(coords=coords, edges=edges, areav=areav, vel=vel, q=z,
                dqdx=np.zeros((n, 3)), rho=rho, visc=visc, mdot=mdot)


# ---------------------------------------------------------------- boxes

def box_axes(nx, ny, nz, stretch_z=1.0):
    """1-D node coordinates of a [0,1]^3 box, geometric stretching of the z
    spacing toward z = 0 by `stretch_z` per cell."""
    xs = np.linspace(0.0, 1.0, nx + 1)
    ys = np.linspace(0.0, 1.0, ny + 1)
    if stretch_z == 1.0:
        zs = np.linspace(0.0, 1.0, nz + 1)
    else:
        h = np.power(stretch_z, np.arange(nz))
        zs = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
    return xs, ys, zs


def hex_box(nx, ny, nz, stretch_z=1.0, jitter=0.0, seed=0):
    """Structured Hex8 box: nodes numbered x fastest, exodus element node
    order, interior nodes moved by +-jitter of the local spacing."""
    xs, ys, zs = box_axes(nx, ny, nz, stretch_z)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    coords = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                       Z.transpose(2, 1, 0).ravel()], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        hx = np.diff(xs).min()
        hy = np.diff(ys).min()
        ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                                 np.arange(nz + 1), indexing="ij")
        ii, jj, kk = (a.transpose(2, 1, 0).ravel() for a in (ii, jj, kk))
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) &
                    (kk < nz))
        dz = np.diff(zs)
        hz = np.minimum(dz[np.clip(kk - 1, 0, nz - 1)], dz[np.clip(kk, 0, nz - 1)])
        d = rng.uniform(-jitter, jitter, size=(coords.shape[0], 3))
        d[:, 0] *= hx
        d[:, 1] *= hy
        d[:, 2] *= hz
        coords = coords + d * interior[:, None]
    px, py = nx + 1, ny + 1
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                          indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    n0 = i + px * (j + py * k)
    conn = np.stack([n0, n0 + 1, n0 + 1 + px, n0 + px, n0 + px * py,
                     n0 + 1 + px * py, n0 + 1 + px + px * py, n0 + px + px * py],
                    axis=1).astype(np.int32)
    return coords, conn

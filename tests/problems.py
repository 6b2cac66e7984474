"""Seeded edge-assembly problems for parity runs: a Hex8 box mesh (product
generator), its geometry, and random node/edge fields, with helpers that
load them into a GPU context and into the oracle."""
import numpy as np

from oracle import orc


def nw():
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = os.path.join(root, "nalu-wind_amd")
    if p not in sys.path:
        sys.path.insert(0, p)
    import naluwind_amd
    return naluwind_amd


class Problem:
    def __init__(self, nx, ny, nz, stretch=1.0, jitter=0.0, seed=0, mixed=None):
        """Hex8 box (nx, ny, nz, stretch, jitter), or with mixed =
        (f_hex, f_tet) an nx^3 conforming mixed Hex8/Tet4/Wed6/Pyr5 box
        (RCM-numbered)."""
        m = nw()
        if mixed is None:
            self.coords, self.conn = m.hex_box(nx, ny, nz, stretch, jitter, seed)
            self.n = self.coords.shape[0]
            self.edges, self.elem_edges = m.hex_edges(self.conn, self.n)
            self.areav = m.hex_edge_area_vector(self.conn, self.coords,
                                                self.elem_edges, self.edges)
            self.dualvol = m.hex_dual_volume(self.conn, self.coords)
        else:
            self.coords, self.types, self.conn = m.mixed_box(
                nx, mixed[0], mixed[1], jitter, seed)
            self.n = self.coords.shape[0]
            self.edges = m.mixed_edges(self.types, self.conn, self.n)
            self.areav = m.mixed_edge_area_vector(self.types, self.conn,
                                                  self.coords, self.edges)
            self.dualvol = np.ones(self.n)  # not read by the edge kernels
        self._random_fields(seed)

    def _random_fields(self, seed):
        m = nw()
        self.ne = self.edges.shape[0]
        rng = np.random.default_rng(seed + 1000)
        n, ne = self.n, self.ne
        self.vel = rng.normal(size=(n, 3))
        self.rho = rng.uniform(0.5, 2.0, n)
        self.p = rng.normal(size=n)
        self.dpdx = rng.normal(size=(n, 3))
        self.udiag = rng.uniform(0.5, 2.0, n)
        self.visc = rng.uniform(1e-3, 1e-1, n)
        self.dudx = rng.normal(size=(n, 9))
        self.wall = (rng.uniform(size=n) < 0.8).astype(float)
        self.q = rng.normal(size=n)
        self.dqdx = rng.normal(size=(n, 3))
        self.dflux = rng.uniform(1e-3, 1e-1, n)
        self.mdot = m.mdot_edge(self.edges, self.areav, self.rho, self.vel)
        self.pecfac = rng.uniform(0.0, 1.0, ne)

    @classmethod
    def from_graph(cls, coords, edges, areav, seed=0):
        """Any edge graph (coords [n][3], edges [e][2], areav [e][3]) with the
        same seeded random fields as the meshes above."""
        P = cls.__new__(cls)
        P.coords = np.ascontiguousarray(coords, dtype=np.float64)
        P.n = P.coords.shape[0]
        P.edges = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 2)
        P.areav = np.ascontiguousarray(areav, dtype=np.float64).reshape(-1, 3)
        P.dualvol = np.ones(P.n)
        P._random_fields(seed)
        return P

    def upload(self, ctx):
        m = nw()
        ctx.mesh_set(self.n, self.edges)
        for fid, a in [(m.F_COORDINATES, self.coords), (m.F_VELOCITY, self.vel),
                       (m.F_DENSITY, self.rho), (m.F_PRESSURE, self.p),
                       (m.F_DPDX, self.dpdx), (m.F_UDIAG, self.udiag),
                       (m.F_VISCOSITY, self.visc), (m.F_DUDX, self.dudx),
                       (m.F_WALL_MASK, self.wall), (m.F_SCALAR_Q, self.q),
                       (m.F_DQDX, self.dqdx), (m.F_DIFF_FLUX_COEFF, self.dflux),
                       (m.F_DUAL_VOLUME, self.dualvol),
                       (m.E_AREA_VECTOR, self.areav),
                       (m.E_MASS_FLOW_RATE, self.mdot),
                       (m.E_PECLET_FACTOR, self.pecfac)]:
            ctx.field_upload(fid, a)

    # oracle assemblies into a fresh CSR
    def oracle_continuity(self, o):
        A = orc.CsrSystem(self.n, self.edges, 1)
        orc.assemble_continuity(o, self.n, self.edges, self.coords, self.vel,
                                self.dpdx, self.rho, self.p, self.udiag,
                                self.areav, csr=A)
        return A

    def oracle_scalar(self, o):
        A = orc.CsrSystem(self.n, self.edges, 1)
        orc.assemble_scalar(o, self.n, self.edges, self.coords, self.vel,
                            self.q, self.dqdx, self.rho, self.dflux, self.areav,
                            self.mdot, csr=A)
        return A

    def oracle_momentum(self, o):
        A = orc.CsrSystem(self.n, self.edges, 3)
        orc.assemble_momentum(o, self.n, self.edges, self.coords, self.vel,
                              self.dudx, self.visc, self.rho, self.areav,
                              self.mdot, self.pecfac, self.wall, csr=A)
        return A


def to_orc_scalar(o):
    return orc.ScalarOpts(o.alpha, o.alpha_upw, o.upw, o.relax, o.use_limiter,
                          orc.Peclet(o.peclet.form, o.peclet.hybrid,
                                     o.peclet.c1, o.peclet.c2))


def to_orc_momentum(o):
    return orc.MomentumOpts(o.alpha, o.alpha_upw, o.upw, o.relax,
                            o.include_divU, o.use_limiter)


def to_orc_continuity(o):
    return orc.ContinuityOpts(o.dt, o.gamma1, o.noc, o.interp_together,
                              o.incompressible)


def assert_csr_close(got, A, rtol=1e-12):
    """got = (row_ptr, cols, vals, rhs) from the GPU; A an oracle CsrSystem.
    Structure must match exactly; values within rtol of the row scale
    max(|a|, max_row |a|) -- the row-scaled 1e-12 parity bar."""
    rp, cols, vals, rhs = got
    np.testing.assert_array_equal(rp, A.row_ptr)
    np.testing.assert_array_equal(cols, A.cols)
    ov = A.vals
    cnt = np.diff(rp)
    rows = np.repeat(np.arange(len(rp) - 1), cnt)
    rowmax = np.zeros(len(rp) - 1)
    full = cnt > 0
    if full.any():
        rowmax[full] = np.maximum.reduceat(np.abs(ov), rp[:-1][full])
    scale = np.maximum(np.abs(ov), rowmax[rows])
    err = np.abs(vals - ov)
    bad = err > rtol * scale + 1e-300
    assert not bad.any(), "max scaled err %.3e" % (err / np.maximum(scale, 1e-300)).max()
    orhs = A.rhs
    rscale = np.maximum(np.abs(orhs), rowmax)
    rerr = np.abs(rhs - orhs)
    assert (rerr <= rtol * rscale + 1e-300).all(), \
        "rhs max scaled err %.3e" % (rerr / np.maximum(rscale, 1e-300)).max()

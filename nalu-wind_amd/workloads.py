"""Synthetic workloads of the edge-assembly / pressure-solve benchmarks.

Each builder returns a Workload: a Hex8 box mesh (library generator), its
edges and geometry, and the seeded node/edge fields of one configuration:

  scalar_tg      uniform box, Taylor-Green velocity u = sin x cos y cos z,
                 v = -cos x sin y cos z, w = 0; rho = 1, mu = 1e-3,
                 phi ~ U(0,1) seeded, dphi/dx from the edge nodal gradient,
                 mdot = 0.5 (rho u_L + rho u_R).A
  momentum_abl   box stretched 1.05 per cell toward z = 0, +-10 % node
                 jitter, log-law u(z) plus 5 % Gaussian noise, nu_t ~
                 U(0, 1e-2), mu_eff = mu + rho nu_t, du/dx from the edge
                 nodal gradient, scheme alpha = 0.5
  continuity_rf  box with the z grading of the momentum box (per-cell ratio
                 1.05^(128/n): the 128-cell geometry refined), divergence-free
                 random Fourier velocity (8 modes), rho = 1, p = 0,
                 dp/dx = 0, Udiag = gamma1/dt; Dirichlet p = 0 on the top
                 face z = 1
  continuity_mixed  the continuity fields on the unstructured mixed mesh: an
                 n^3 box cut into Hex8/Tet4/Wed6 cells (60/30/10 % of the
                 cells, seeded), conforming through Pyr5 + Tet4 transition
                 cells about a centre node, +-10 % node jitter, nodes
                 randomly permuted and then RCM-renumbered by the library's
                 generator
"""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import naluwind_amd as nw  # noqa: E402


class Workload:
    def __init__(self, name, nx, ny, nz, stretch, jitter, seed):
        self.name = name
        self.cells = (nx, ny, nz)
        self.coords, self.conn = nw.hex_box(nx, ny, nz, stretch, jitter, seed)
        self.n = self.coords.shape[0]
        self.edges, elem_edges = nw.hex_edges(self.conn, self.n)
        self.ne = self.edges.shape[0]
        self.areav = nw.hex_edge_area_vector(self.conn, self.coords, elem_edges,
                                             self.edges)
        del elem_edges
        self.dualvol = nw.hex_dual_volume(self.conn, self.coords)
        self.fields = {}   # field id -> array
        self.top = None    # Dirichlet nodes (continuity)

    def upload(self, ctx):
        ctx.mesh_set(self.n, self.edges)
        ctx.field_upload(nw.F_COORDINATES, self.coords)
        ctx.field_upload(nw.E_AREA_VECTOR, self.areav)
        if self.dualvol is not None:
            ctx.field_upload(nw.F_DUAL_VOLUME, self.dualvol)
        for fid, a in self.fields.items():
            ctx.field_upload(fid, a)

    def drop_elements(self):
        self.conn = None


def scalar_tg(n=32, seed=1234):
    W = Workload("scalar_tg", n, n, n, 1.0, 0.0, seed)
    x, y, z = W.coords.T
    vel = np.stack([np.sin(x) * np.cos(y) * np.cos(z),
                    -np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(z)], 1)
    rho = np.ones(W.n)
    phi = np.random.default_rng(seed).uniform(0.0, 1.0, W.n)
    W.fields = {
        nw.F_VELOCITY: vel, nw.F_DENSITY: rho, nw.F_SCALAR_Q: phi,
        nw.F_DIFF_FLUX_COEFF: np.full(W.n, 1.0e-3),
        nw.F_DQDX: nw.nodal_grad_edge(W.edges, W.areav, W.dualvol, phi),
        nw.E_MASS_FLOW_RATE: nw.mdot_edge(W.edges, W.areav, rho, vel),
    }
    W.opts = nw.scalar_defaults()
    W.drop_elements()
    return W


def momentum_abl(n=128, seed=1234, mu=1.8e-5):
    W = Workload("momentum_abl", n, n, n, 1.05, 0.10, seed)
    rng = np.random.default_rng(seed)
    z = W.coords[:, 2]
    uref, z0 = 10.0, 1.0e-3
    vel = np.zeros((W.n, 3))
    vel[:, 0] = uref * np.log1p(z / z0) / math.log1p(1.0 / z0)
    vel += 0.05 * uref * rng.standard_normal((W.n, 3))
    rho = np.ones(W.n)
    nut = rng.uniform(0.0, 1.0e-2, W.n)
    W.fields = {
        nw.F_VELOCITY: vel, nw.F_DENSITY: rho, nw.F_VISCOSITY: mu + rho * nut,
        nw.F_DUDX: nw.nodal_grad_edge(W.edges, W.areav, W.dualvol, vel),
        nw.F_WALL_MASK: np.ones(W.n),
        nw.E_MASS_FLOW_RATE: nw.mdot_edge(W.edges, W.areav, rho, vel),
    }
    W.opts = nw.momentum_defaults(alpha=0.5)
    W.peclet = nw.Peclet(0, 1.0, 2.0, 4.0)   # hybrid factor 1: blend active
    W.drop_elements()
    return W


def fourier_modes(seed, modes=8):
    """Seeded modes (k [m][3], b [m][3], phase [m], omega [m]) with b_m
    orthogonal to k_m, so that each mode is divergence-free."""
    rng = np.random.default_rng(seed)
    K, B, PH, OM = [], [], [], []
    for _ in range(modes):
        k = 2.0 * math.pi * rng.integers(1, 4, size=3) * rng.choice([-1, 1], 3)
        b = rng.standard_normal(3)
        b -= (b @ k) / (k @ k) * k
        K.append(k)
        B.append(b)
        PH.append(rng.uniform(0, 2 * math.pi))
        OM.append(rng.uniform(0.5, 2.0))
    return (np.array(K, float), np.array(B, float), np.array(PH), np.array(OM))


def fourier_velocity(coords, seed, modes=8, t=0.0):
    """Divergence-free sum of 8 seeded Fourier modes b_m sin(k_m.x + phi_m +
    w_m t) with b_m orthogonal to k_m."""
    K, B, PH, OM = fourier_modes(seed, modes)
    u = np.zeros_like(coords)
    for k, b, ph, om in zip(K, B, PH, OM):
        u += np.sin(coords @ k + ph + om * t)[:, None] * b[None, :]
    return u


def grading_128(n):
    """Per-cell z ratio giving an n-cell box the grading of the 128-cell
    momentum box (1.05 per cell, first/last spacing ratio 1.05^127): the
    same stretched geometry at another resolution."""
    return 1.05 ** (128.0 / n)


def continuity_rf(n=256, seed=1234, stretch=None, dt=1.0e-3, t=0.0,
                  keep_elements=False):
    if stretch is None:
        stretch = grading_128(n)
    W = Workload("continuity_rf", n, n, n, stretch, 0.0, seed)
    vel = fourier_velocity(W.coords, seed, t=t)
    W.fields = {
        nw.F_VELOCITY: vel, nw.F_DENSITY: np.ones(W.n),
        nw.F_PRESSURE: np.zeros(W.n), nw.F_DPDX: np.zeros((W.n, 3)),
        nw.F_UDIAG: np.full(W.n, 1.0 / dt),
        nw.F_PRESSURE_BC: np.zeros(W.n),
    }
    W.opts = nw.continuity_defaults(dt=dt, gamma1=1.0)
    W.top = np.where(W.coords[:, 2] >= 1.0 - 1e-12)[0].astype(np.int32)
    W.seed = seed
    if not keep_elements:
        W.drop_elements()
    return W


class MixedWorkload(Workload):
    """Conforming mixed Hex8/Tet4/Wed6/Pyr5 box from the library's
    generator (RCM node order), its edges and edge area vectors."""

    def __init__(self, name, n, f_hex, f_tet, jitter, seed, keep_elements=False):
        self.name = name
        self.cells = (n, n, n)
        self.coords, types, conn = nw.mixed_box(n, f_hex, f_tet, jitter, seed)
        self.n = self.coords.shape[0]
        self.edges = nw.mixed_edges(types, conn, self.n)
        self.ne = self.edges.shape[0]
        self.areav = nw.mixed_edge_area_vector(types, conn, self.coords,
                                               self.edges)
        self.elements = np.bincount(types, minlength=4)  # hex, tet, wedge, pyramid
        self.dualvol = None
        self.fields = {}
        self.top = None
        self.types, self.conn = (types, conn) if keep_elements else (None, None)


def continuity_mixed(n=160, seed=1234, dt=1.0e-3, f_hex=0.6, f_tet=0.3,
                     jitter=0.10, keep_elements=False):
    W = MixedWorkload("continuity_mixed", n, f_hex, f_tet, jitter, seed,
                      keep_elements)
    vel = fourier_velocity(W.coords, seed)
    W.fields = {
        nw.F_VELOCITY: vel, nw.F_DENSITY: np.ones(W.n),
        nw.F_PRESSURE: np.zeros(W.n), nw.F_DPDX: np.zeros((W.n, 3)),
        nw.F_UDIAG: np.full(W.n, 1.0 / dt),
        nw.F_PRESSURE_BC: np.zeros(W.n),
    }
    W.opts = nw.continuity_defaults(dt=dt, gamma1=1.0)
    W.top = np.where(W.coords[:, 2] >= 1.0 - 1e-12)[0].astype(np.int32)
    W.seed = seed
    return W


def apply_pressure_bc(ctx, W):
    """Dirichlet p = 0 on the top face: rows of the top nodes become the
    identity with rhs = pressure_bc - p (pressure_bc = 0)."""
    ctx.apply_dirichlet(W.top, nw.F_PRESSURE, nw.F_PRESSURE_BC, 0, 1)


class PressureSeries:
    """The pressure-projection step of the low-Mach algorithm
    (LowMachEquationSystem::solve_and_update, src/LowMachEquationSystem.C:
    793-819) on the continuity_rf box, one outer iteration per time step:

      step(k)   the momentum predictor: the velocity of the seeded random
                Fourier flow at t = k dt_flow (generated on the device);
                then the continuity assembly with the current pressure and
                pressure gradient, and Dirichlet rows on the top face
                (dp = pressure_bc - p = 0 there);
      the caller solves for the pressure increment dp;
      update()  p += dp (solution_update), mdot from the new pressure
                (MdotEdgeAlg) and the nodal velocity projection, which
                recomputes dp/dx from the new pressure.

    p and dp/dx start at zero, so step 0 solves for the whole pressure and
    later steps for the corrections the changing flow and the updated
    pressure gradient ask for.  The matrix is the same at every step (a
    static mesh with Udiag = gamma1/dt); the right-hand side is not."""

    def __init__(self, ctx, n=256, seed=1234, stretch=None, dt_flow=0.02, W=None):
        """A new continuity_rf(n, seed, stretch) box loaded into ctx, or a
        continuity_rf workload W that ctx already holds with its 1-dof
        graph (its pressure and pressure gradient are reset to zero)."""
        self.ctx = ctx
        if W is None:
            W = continuity_rf(n, seed, stretch)
            W.upload(ctx)
            ctx.build_graph(1)
        else:
            ctx.field_fill(nw.F_PRESSURE, 0.0)
            ctx.field_fill(nw.F_DPDX, 0.0)
        self.W = W
        self.dt_flow = dt_flow
        self.seed = W.seed
        self.modes = fourier_modes(W.seed)

    def step(self, k):
        """Predictor velocity at t = k dt_flow, continuity assembly with the
        current p and dp/dx, Dirichlet rows."""
        self.ctx.fourier_velocity(self.modes, k * self.dt_flow)
        self.ctx.assemble_continuity(self.W.opts)
        apply_pressure_bc(self.ctx, self.W)

    def update(self):
        """After the solve: p += dp, mdot, velocity projection (dp/dx)."""
        self.ctx.solution_update(1.0, nw.LINSYS_SOLUTION, 1.0, nw.F_PRESSURE)
        self.ctx.compute_mdot(1.0, 1.0)
        self.ctx.project_nodal_velocity()

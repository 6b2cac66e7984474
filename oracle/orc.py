"""ctypes wrapper of the C oracle (oracle/build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product library
never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liborc.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


lib = _load()


class Peclet(C.Structure):
    _fields_ = [("form", C.c_int), ("hybrid", C.c_double), ("c1", C.c_double),
                ("c2", C.c_double)]


class ScalarOpts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("alpha_upw", C.c_double),
                ("upw", C.c_double), ("relax", C.c_double),
                ("use_limiter", C.c_int), ("peclet", Peclet)]


class MomentumOpts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("alpha_upw", C.c_double),
                ("upw", C.c_double), ("relax", C.c_double),
                ("include_divU", C.c_double), ("use_limiter", C.c_int)]


class Buoyancy(C.Structure):
    """orc_buoyancy: gravity, buoyancy_source [n][3], buoyancy_source_mask."""
    _fields_ = [("gravity", C.c_double * 3), ("source", C.POINTER(C.c_double)),
                ("mask", C.POINTER(C.c_double))]


class ContinuityOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("gamma1", C.c_double), ("noc", C.c_double),
                ("interp_together", C.c_double),
                ("incompressible", C.c_double),
                ("buoy", C.POINTER(Buoyancy))]


class Csr(C.Structure):
    _fields_ = [("nnode", C.c_int64), ("ndof", C.c_int), ("nrows", C.c_int64),
                ("nnz", C.c_int64), ("row_ptr", C.POINTER(C.c_int64)),
                ("cols", C.POINTER(C.c_int32)), ("vals", C.POINTER(C.c_double)),
                ("rhs", C.POINTER(C.c_double)), ("skip", C.c_void_p),
                ("skip_check", C.c_int)]


class GmresOpts(C.Structure):
    _fields_ = [("kspace", C.c_int), ("max_iters", C.c_int), ("tol", C.c_double),
                ("precond", C.c_int), ("cheb_degree", C.c_int),
                ("cheb_ratio", C.c_double)]


class GmresStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("implicit_rel", C.c_double),
                ("final_abs", C.c_double), ("bnorm", C.c_double),
                ("converged", C.c_int)]


PREC_NONE, PREC_JACOBI, PREC_CHEBYSHEV = 0, 1, 2

lib.orc_hex_scs_areav.argtypes = [_f64p, _f64p]
lib.orc_hex_scv_volume.argtypes = [_f64p, _f64p]
lib.orc_create_edges_hex.argtypes = [C.c_int64, _i32p, C.c_int64, _i32p,
                                     C.c_int64, _i32p]
lib.orc_create_edges_hex.restype = C.c_int64
lib.orc_edge_area_vec_hex.argtypes = [C.c_int64, _i32p, _f64p, _i32p, _i32p,
                                      C.c_int64, _f64p]
lib.orc_dual_volume_hex.argtypes = [C.c_int64, _i32p, _f64p, C.c_int64, _f64p]
lib.orc_nodal_grad_edge.argtypes = [C.c_int64, C.c_int64, _i32p, _f64p, _f64p,
                                    C.c_int, _f64p, _f64p]
lib.orc_mdot_edge.argtypes = [C.c_int64, _i32p, _f64p, _f64p, _f64p, _f64p]
lib.orc_mdot_edge_alg.argtypes = [C.c_int64, _i32p, _f64p, _f64p, _f64p, _f64p,
                                  _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                  C.POINTER(Buoyancy), _f64p]
lib.orc_peclet_classic.argtypes = [C.c_double, C.c_double]
lib.orc_peclet_classic.restype = C.c_double
lib.orc_peclet_tanh.argtypes = [C.c_double, C.c_double, C.c_double]
lib.orc_peclet_tanh.restype = C.c_double
lib.orc_van_leer.argtypes = [C.c_double, C.c_double, C.c_double]
lib.orc_van_leer.restype = C.c_double
lib.orc_momentum_edge_peclet.argtypes = [C.c_int64, _i32p, _f64p, _f64p, _f64p,
                                         _f64p, C.POINTER(Peclet), C.c_void_p,
                                         _f64p]
lib.orc_scalar_edge.argtypes = [C.POINTER(ScalarOpts), C.c_int64, C.c_int32,
                                C.c_int32] + [_f64p] * 8 + [_f64p, _f64p]
lib.orc_momentum_edge.argtypes = [C.POINTER(MomentumOpts), C.c_int64,
                                  C.c_int32, C.c_int32] + [_f64p] * 9 + [_f64p,
                                                                         _f64p]
lib.orc_continuity_edge.argtypes = [C.POINTER(ContinuityOpts), C.c_int64,
                                    C.c_int32, C.c_int32] + [_f64p] * 7 + [
                                        _f64p, _f64p]
lib.orc_csr_build_edge_graph.argtypes = [C.c_int64, C.c_int64, _i32p, C.c_int,
                                         C.POINTER(Csr)]
lib.orc_csr_free.argtypes = [C.POINTER(Csr)]
lib.orc_csr_sum_into.argtypes = [C.POINTER(Csr), C.c_int, _i32p, _f64p, _f64p]
lib.orc_csr_reset_rows.argtypes = [C.POINTER(Csr), C.c_int64, _i32p, C.c_int,
                                   C.c_int, C.c_double, C.c_double]
lib.orc_csr_apply_dirichlet.argtypes = [C.POINTER(Csr), C.c_int64, _i32p,
                                        C.c_int, C.c_int, _f64p, _f64p, C.c_int]
lib.orc_csr_fix_node.argtypes = [C.POINTER(Csr), C.c_int32, C.c_double,
                                 C.c_double]
lib.orc_csr_zero.argtypes = [C.POINTER(Csr)]
lib.orc_csr_apply_dirichlet_hypre.argtypes = [C.POINTER(Csr), C.c_int64, _i32p,
                                              _f64p, _f64p, C.c_int]
_dense = C.c_void_p
lib.orc_assemble_scalar.argtypes = [C.POINTER(ScalarOpts), C.c_int64,
                                    C.c_int64, _i32p] + [_f64p] * 8 + [
                                        C.c_void_p, _dense, _dense]
lib.orc_assemble_momentum.argtypes = [C.POINTER(MomentumOpts), C.c_int64,
                                      C.c_int64, _i32p] + [_f64p] * 9 + [
                                          C.c_void_p, _dense, _dense]
lib.orc_assemble_continuity.argtypes = [C.POINTER(ContinuityOpts), C.c_int64,
                                        C.c_int64, _i32p] + [_f64p] * 7 + [
                                            C.c_void_p, _dense, _dense]
lib.orc_set_threads.argtypes = [C.c_int]
lib.orc_tet_scs_areav.argtypes = [_f64p, _f64p]
lib.orc_wed_scs_areav.argtypes = [_f64p, _f64p]
lib.orc_edge_area_vec_mixed.argtypes = [C.c_int64, _i32p, _i32p, _f64p, _i32p,
                                        C.c_int64, _f64p]
lib.orc_edge_area_vec_mixed.restype = C.c_int
lib.orc_pyr_scs_areav.argtypes = [_f64p, _f64p]
lib.orc_pyr_scv_volume.argtypes = [_f64p, _f64p]
lib.orc_tet_scv_volume.argtypes = [_f64p, _f64p]
lib.orc_wed_scv_volume.argtypes = [_f64p, _f64p]
lib.orc_dual_volume_mixed.argtypes = [C.c_int64, _i32p, _i32p, _f64p, C.c_int64,
                                      _f64p]
lib.orc_spmv.argtypes = [C.POINTER(Csr), _f64p, _f64p]
lib.orc_gmres.argtypes = [C.POINTER(Csr), _f64p, _f64p, C.POINTER(GmresOpts),
                          C.POINTER(GmresStats)]
lib.orc_gmres.restype = C.c_int
lib.orc_dense_solve.argtypes = [C.c_int, _f64p, _f64p]
lib.orc_dense_solve.restype = C.c_int
lib.orc_jacobi_lambda_max.argtypes = [C.POINTER(Csr), C.c_int, C.c_uint64]
lib.orc_jacobi_lambda_max.restype = C.c_double


def set_threads(n):
    """Threads of the CSR assembly loops (atomic scatter above 1)."""
    lib.orc_set_threads(int(n))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- geometry

def hex_scs_areav(coords8):
    out = np.zeros(36)
    lib.orc_hex_scs_areav(f64(coords8).ravel(), out)
    return out.reshape(12, 3)


def hex_scv_volume(coords8):
    out = np.zeros(8)
    lib.orc_hex_scv_volume(f64(coords8).ravel(), out)
    return out


def tet_scs_areav(coords4):
    out = np.zeros(18)
    lib.orc_tet_scs_areav(f64(coords4).ravel(), out)
    return out.reshape(6, 3)


def wed_scs_areav(coords6):
    out = np.zeros(27)
    lib.orc_wed_scs_areav(f64(coords6).ravel(), out)
    return out.reshape(9, 3)


TET_LRSCV = np.array([0, 1, 1, 2, 0, 2, 0, 3, 1, 3, 2, 3]).reshape(6, 2)
WED_LRSCV = np.array([0, 1, 1, 2, 0, 2, 3, 4, 4, 5, 3, 5, 0, 3, 1, 4, 2,
                      5]).reshape(9, 2)
HEX_LRSCV = np.array([0, 1, 1, 2, 2, 3, 0, 3, 4, 5, 5, 6, 6, 7, 4, 7, 0, 4,
                      1, 5, 2, 6, 3, 7]).reshape(12, 2)


def tet_scv_volume(coords4):
    out = np.zeros(4)
    lib.orc_tet_scv_volume(f64(coords4).ravel(), out)
    return out


def wed_scv_volume(coords6):
    out = np.zeros(6)
    lib.orc_wed_scv_volume(f64(coords6).ravel(), out)
    return out


def pyr_scs_areav(coords5):
    out = np.zeros(36)
    lib.orc_pyr_scs_areav(f64(coords5).ravel(), out)
    return out.reshape(12, 3)


def pyr_scv_volume(coords5):
    out = np.zeros(5)
    lib.orc_pyr_scv_volume(f64(coords5).ravel(), out)
    return out


PYR_LRSCV = np.array([0, 1, 1, 2, 2, 3, 0, 3, 0, 4, 0, 4, 1, 4, 1, 4, 2, 4,
                      2, 4, 3, 4, 3, 4]).reshape(12, 2)


def dual_volume_mixed(types, conn, coords):
    coords = f64(coords)
    out = np.zeros(coords.shape[0])
    lib.orc_dual_volume_mixed(len(types), i32(types), i32(conn).ravel(),
                              coords.ravel(), coords.shape[0], out)
    return out


def edge_area_vec_mixed(types, conn, coords, edges):
    out = np.zeros((edges.shape[0], 3))
    rc = lib.orc_edge_area_vec_mixed(len(types), i32(types), i32(conn).ravel(),
                                     f64(coords).ravel(), i32(edges).ravel(),
                                     edges.shape[0], out)
    assert rc == 0
    return out


def create_edges_hex(conn, nnode):
    conn = i32(conn)
    nelem = conn.size // 8
    cap = 3 * nnode + 12 * nelem
    edges = np.zeros(2 * cap, np.int32)
    elem_edges = np.zeros(12 * nelem, np.int32)
    ne = lib.orc_create_edges_hex(nelem, conn.ravel(), nnode, edges, cap,
                                  elem_edges)
    assert ne >= 0
    return edges[:2 * ne].reshape(ne, 2).copy(), elem_edges.reshape(nelem, 12)


def edge_area_vec_hex(conn, coords, elem_edges, edges):
    conn = i32(conn)
    out = np.zeros((edges.shape[0], 3))
    lib.orc_edge_area_vec_hex(conn.size // 8, conn.ravel(), f64(coords).ravel(),
                              i32(elem_edges).ravel(), i32(edges).ravel(),
                              edges.shape[0], out)
    return out


def dual_volume_hex(conn, coords):
    conn = i32(conn)
    coords = f64(coords)
    out = np.zeros(coords.shape[0])
    lib.orc_dual_volume_hex(conn.size // 8, conn.ravel(), coords.ravel(),
                            coords.shape[0], out)
    return out


def nodal_grad_edge(edges, areav, dualvol, phi):
    phi = f64(phi)
    n = dualvol.shape[0]
    dim1 = 1 if phi.ndim == 1 else phi.shape[1]
    out = np.zeros((n, dim1 * 3))
    lib.orc_nodal_grad_edge(n, edges.shape[0], i32(edges).ravel(),
                            f64(areav).ravel(), f64(dualvol), dim1,
                            phi.ravel(), out)
    return out


def mdot_edge(edges, areav, rho, vel):
    out = np.zeros(edges.shape[0])
    lib.orc_mdot_edge(edges.shape[0], i32(edges).ravel(), f64(areav).ravel(),
                      f64(rho), f64(vel).ravel(), out)
    return out


def buoyancy(gravity, source, mask):
    """The balanced buoyancy force inputs (orc_buoyancy): gravity vector,
    buoyancy_source [n][3], buoyancy_source_mask [n]; the arrays are kept
    alive by the returned structure."""
    src = np.ascontiguousarray(source, dtype=np.float64).ravel()
    msk = np.ascontiguousarray(mask, dtype=np.float64).ravel()
    b = Buoyancy((C.c_double * 3)(*[float(g) for g in gravity]),
                 src.ctypes.data_as(C.POINTER(C.c_double)),
                 msk.ctypes.data_as(C.POINTER(C.c_double)))
    b.keep_alive = (src, msk)
    return b


def with_buoyancy(opts, buoy):
    """A copy of ContinuityOpts with the balanced buoyancy force on."""
    o = ContinuityOpts(opts.dt, opts.gamma1, opts.noc, opts.interp_together,
                       opts.incompressible, C.pointer(buoy))
    return o


def mdot_edge_alg(edges, coords, vel, dpdx, rho, p, udiag, areav, noc=1.0,
                  interp=1.0, buoy=None):
    """MdotEdgeAlg interior edges: node arrays [n][3] / [n], areav [e][3];
    buoy: a `buoyancy(...)` structure for the balanced buoyancy force."""
    out = np.zeros(edges.shape[0])
    lib.orc_mdot_edge_alg(edges.shape[0], i32(edges).ravel(), f64(coords).ravel(),
                          f64(vel).ravel(), f64(dpdx).ravel(), f64(rho), f64(p),
                          f64(udiag), f64(areav).ravel(), noc, interp,
                          C.byref(buoy) if buoy is not None else None, out)
    return out


def solution_update(delta_frac, delta, field_frac, field):
    """EquationSystem::solution_update (src/EquationSystem.C:556-573,
    nalu_ngp::field_axpby): field_frac * field + delta_frac * delta."""
    return delta_frac * f64(delta) + field_frac * f64(field)


def project_nodal_velocity(edges, areav, dualvol, p, dpdx_old, vel, rho, udiag):
    """LowMachEquationSystem::project_nodal_velocity (src/LowMachEquation
    System.C:882-939): dp/dx of the new pressure by the edge nodal gradient
    (:913), then u -= (dp/dx_new - dp/dx_old) / (rho Udiag) at every node
    (:930-939).  Returns (dpdx_new, vel_new)."""
    g = nodal_grad_edge(edges, areav, dualvol, p)
    fac = 1.0 / (f64(rho) * f64(udiag))
    return g, f64(vel) - fac[:, None] * (g - f64(dpdx_old))


def momentum_edge_peclet(edges, coords, vel, rho, visc, peclet,
                         with_number=False):
    """MomentumEdgePecletAlg: the edge Peclet factor, and with with_number
    the pair (Peclet number, factor)."""
    out = np.zeros(edges.shape[0])
    num = np.zeros(edges.shape[0]) if with_number else None
    lib.orc_momentum_edge_peclet(edges.shape[0], i32(edges).ravel(),
                                 f64(coords).ravel(), f64(vel).ravel(),
                                 f64(rho), f64(visc), C.byref(peclet),
                                 _ptr(num), out)
    return (num, out) if with_number else out


# ---------------------------------------------------------------- CSR

class CsrSystem:
    """Owns an orc_csr built from an edge graph."""

    def __init__(self, nnode, edges, ndof):
        self.c = Csr()
        rc = lib.orc_csr_build_edge_graph(nnode, edges.shape[0],
                                          i32(edges).ravel(), ndof,
                                          C.byref(self.c))
        assert rc == 0

    @classmethod
    def from_dense(cls, D, rhs):
        """1-dof system whose graph is the symmetric nonzero pattern of D."""
        D = f64(D)
        iu, ju = np.nonzero(np.triu((D != 0) | (D.T != 0), 1))
        edges = np.stack([iu, ju], axis=1).astype(np.int32)
        A = cls(D.shape[0], edges, 1)
        rp, cols, vals = A.row_ptr, A.cols, A.vals
        for r in range(D.shape[0]):
            for k in range(rp[r], rp[r + 1]):
                vals[k] = D[r, cols[k]]
        A.rhs[:] = f64(rhs)
        return A

    def __del__(self):
        try:
            lib.orc_csr_free(C.byref(self.c))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self.c)

    def _view(self, ptr, n, dt):
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dt)

    @property
    def row_ptr(self):
        return np.ctypeslib.as_array(self.c.row_ptr, shape=(self.c.nrows + 1,))

    @property
    def cols(self):
        return np.ctypeslib.as_array(self.c.cols, shape=(self.c.nnz,))

    @property
    def vals(self):
        return np.ctypeslib.as_array(self.c.vals, shape=(self.c.nnz,))

    @property
    def rhs(self):
        return np.ctypeslib.as_array(self.c.rhs, shape=(self.c.nrows,))

    def dense(self):
        n = self.c.nrows
        D = np.zeros((n, n))
        rp, cols, vals = self.row_ptr, self.cols, self.vals
        for r in range(n):
            for k in range(rp[r], rp[r + 1]):
                D[r, cols[k]] += vals[k]
        return D

    def to_scipy(self):
        import scipy.sparse as sp
        n = self.c.nrows
        return sp.csr_matrix((self.vals.copy(), self.cols.copy(),
                              self.row_ptr.copy()), shape=(n, n))

    def sum_into(self, nodes, lhs, rhs):
        lib.orc_csr_sum_into(self.ref, len(nodes), i32(nodes), f64(lhs).ravel(),
                             f64(rhs))

    def reset_rows(self, nodes, begin, end, diag=0.0, rhs_res=0.0):
        lib.orc_csr_reset_rows(self.ref, len(nodes), i32(nodes), begin, end,
                               diag, rhs_res)

    def set_skipped_rows(self, nodes):
        """Skipped rows of the hypre backend (orc_csr.skip)."""
        self._skip = np.zeros(self.c.nnode, np.uint8)
        self._skip[np.asarray(nodes, np.int64)] = 1
        self.c.skip = self._skip.ctypes.data_as(C.c_void_p)
        self.c.skip_check = 1

    def zero(self):
        lib.orc_csr_zero(self.ref)

    def apply_dirichlet_hypre(self, nodes, sol, bc):
        sol = f64(sol)
        ncomp = 1 if sol.ndim == 1 else sol.shape[1]
        lib.orc_csr_apply_dirichlet_hypre(self.ref, len(nodes), i32(nodes),
                                          sol.ravel(), f64(bc).ravel(), ncomp)

    def apply_dirichlet(self, nodes, begin, end, sol, bc):
        sol = f64(sol)
        ncomp = 1 if sol.ndim == 1 else sol.shape[1]
        lib.orc_csr_apply_dirichlet(self.ref, len(nodes), i32(nodes), begin,
                                    end, sol.ravel(), f64(bc).ravel(), ncomp)

    def fix_node(self, node, pref, pnode):
        lib.orc_csr_fix_node(self.ref, int(node), pref, pnode)

    def spmv(self, x):
        y = np.zeros(self.c.nrows)
        lib.orc_spmv(self.ref, f64(x), y)
        return y

    def gmres(self, b, x0=None, kspace=50, max_iters=1000, tol=1e-8,
              precond=PREC_JACOBI, cheb_degree=4, cheb_ratio=30.0):
        x = np.zeros(self.c.nrows) if x0 is None else f64(x0).copy()
        o = GmresOpts(kspace, max_iters, tol, precond, cheb_degree, cheb_ratio)
        st = GmresStats()
        lib.orc_gmres(self.ref, f64(b), x, C.byref(o), C.byref(st))
        return x, st

    def jacobi_lambda_max(self, iters=30, seed=12345):
        return lib.orc_jacobi_lambda_max(self.ref, iters, seed)


def dense_solve(A, b):
    A = f64(A).copy()
    x = f64(b).copy()
    rc = lib.orc_dense_solve(A.shape[0], A.ravel(), x)
    assert rc == 0
    return x


# ---------------------------------------------------------------- assembly

def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def assemble_scalar(opts, nnode, edges, coords, vrtm, q, dqdx, rho, dflux,
                    areav, mdot, csr=None):
    """Returns csr (if given) or a dense (lhs, rhs) from the unit-test mock."""
    ld = rd = None
    if csr is None:
        ld = np.zeros((nnode, nnode))
        rd = np.zeros(nnode)
    lib.orc_assemble_scalar(C.byref(opts), nnode, edges.shape[0],
                            i32(edges).ravel(), f64(coords).ravel(),
                            f64(vrtm).ravel(), f64(q), f64(dqdx).ravel(),
                            f64(rho), f64(dflux), f64(areav).ravel(), f64(mdot),
                            csr.ref if csr is not None else None, _ptr(ld),
                            _ptr(rd))
    return csr if csr is not None else (ld, rd)


def assemble_momentum(opts, nnode, edges, coords, vel, dudx, visc, rho, areav,
                      mdot, pecfac, wallmask, csr=None):
    ld = rd = None
    if csr is None:
        ld = np.zeros((3 * nnode, 3 * nnode))
        rd = np.zeros(3 * nnode)
    lib.orc_assemble_momentum(C.byref(opts), nnode, edges.shape[0],
                              i32(edges).ravel(), f64(coords).ravel(),
                              f64(vel).ravel(), f64(dudx).ravel(), f64(visc),
                              f64(rho), f64(areav).ravel(), f64(mdot),
                              f64(pecfac), f64(wallmask),
                              csr.ref if csr is not None else None, _ptr(ld),
                              _ptr(rd))
    return csr if csr is not None else (ld, rd)


def assemble_continuity(opts, nnode, edges, coords, vel, dpdx, rho, p, udiag,
                        areav, csr=None):
    ld = rd = None
    if csr is None:
        ld = np.zeros((nnode, nnode))
        rd = np.zeros(nnode)
    lib.orc_assemble_continuity(C.byref(opts), nnode, edges.shape[0],
                                i32(edges).ravel(), f64(coords).ravel(),
                                f64(vel).ravel(), f64(dpdx).ravel(), f64(rho),
                                f64(p), f64(udiag), f64(areav).ravel(),
                                csr.ref if csr is not None else None, _ptr(ld),
                                _ptr(rd))
    return csr if csr is not None else (ld, rd)


lib.orc_mpnn_nparams.argtypes = [C.c_int, C.c_int]
lib.orc_mpnn_nparams.restype = C.c_int64
lib.orc_mpnn_slot_weights.argtypes = [C.POINTER(Csr), _f64p]
lib.orc_mpnn_features.argtypes = [C.POINTER(Csr), _f64p,
                                  C.POINTER(C.POINTER(C.c_double)), C.c_int,
                                  _f64p, _f64p]
lib.orc_mpnn_base_guess.argtypes = [C.c_int64, C.POINTER(C.POINTER(C.c_double)),
                                    C.c_int, _f64p]
lib.orc_mpnn_forward.argtypes = [_f64p, C.c_int, C.c_int, C.POINTER(Csr), _f64p,
                                 _f64p, _f64p, C.c_void_p]
lib.orc_mpnn_forward_bf16.argtypes = [_f64p, C.c_int, C.c_int, C.POINTER(Csr),
                                      _f64p, _f64p, _f64p]
lib.orc_mpnn_forward_bf16.restype = None
lib.orc_mpnn_loss_grad.argtypes = [_f64p, C.c_int, C.c_int, C.POINTER(Csr),
                                   _f64p, _f64p, _f64p, _f64p]
lib.orc_mpnn_loss_grad.restype = C.c_double
lib.orc_mpnn_loss_grad_residual.argtypes = [_f64p, C.c_int, C.c_int,
                                            C.POINTER(Csr), _f64p, _f64p,
                                            _f64p, _f64p]
lib.orc_mpnn_loss_grad_residual.restype = C.c_double


def mpnn_nparams(H, L):
    return lib.orc_mpnn_nparams(H, L)


def mpnn_slot_weights(A):
    w = np.zeros(A.c.nnz)
    lib.orc_mpnn_slot_weights(A.ref, w)
    return w


MPNN_NF = 5


def _hist_ptrs(hist):
    arrs = [np.ascontiguousarray(h, dtype=np.float64) for h in hist]
    ptrs = (C.POINTER(C.c_double) * len(arrs))(
        *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return arrs, ptrs


def mpnn_features(A, b, *hist):
    """Node features and scales {s_b, s_x, s_r, s_e, s_o} from the history
    of guesses hist = (x0, xp, xpp, ...), newest first, 1 to 5 of them."""
    hist = [h for h in hist if h is not None]
    assert 1 <= len(hist) <= 5
    feat = np.zeros((A.c.nrows, MPNN_NF))
    sc = np.zeros(5)
    arrs, ptrs = _hist_ptrs(hist)
    lib.orc_mpnn_features(A.ref, f64(b), ptrs, len(arrs), feat, sc)
    return feat, sc


def mpnn_base_guess(*hist):
    """The base guess xb: polynomial extrapolation in time of the newest
    min(len(hist), 3) guesses."""
    hist = [h for h in hist if h is not None]
    arrs, ptrs = _hist_ptrs(hist)
    out = np.zeros(arrs[0].shape[0])
    lib.orc_mpnn_base_guess(out.shape[0], ptrs, len(arrs), out)
    return out


def mpnn_forward(P, H, L, A, w, feat):
    d = np.zeros(A.c.nrows)
    lib.orc_mpnn_forward(f64(P), H, L, A.ref, f64(w), f64(feat), d, None)
    return d


def mpnn_forward_bf16(P, H, L, A, w, feat):
    d = np.zeros(A.c.nrows)
    lib.orc_mpnn_forward_bf16(f64(P), H, L, A.ref, f64(w), f64(feat), d)
    return d


def mpnn_loss_grad(P, H, L, A, w, feat, target):
    g = np.zeros(len(P))
    loss = lib.orc_mpnn_loss_grad(f64(P), H, L, A.ref, f64(w), f64(feat),
                                  f64(target), g)
    return loss, g


def mpnn_loss_grad_residual(P, H, L, A, w, feat, target):
    g = np.zeros(len(P))
    loss = lib.orc_mpnn_loss_grad_residual(f64(P), H, L, A.ref, f64(w),
                                           f64(feat), f64(target), g)
    return loss, g


def peclet_classic(hf, pe):
    return lib.orc_peclet_classic(hf, pe)


def peclet_tanh(c1, c2, x):
    return lib.orc_peclet_tanh(c1, c2, x)


def van_leer(a, b, eps=1e-16):
    return lib.orc_van_leer(a, b, eps)

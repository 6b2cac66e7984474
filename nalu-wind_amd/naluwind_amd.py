"""Python binding (ctypes) of libnaluwind_amd.so, the MI355X edge-assembly +
pressure-solve library.  Mirrors the C ABI in include/nalu_amd.h and
include/nalu_amd_mesh.h one to one; numpy arrays in, numpy arrays out.

The library must be built (nalu-wind_amd/build.py).  Importing this module
without it raises ImportError: there is no fallback implementation.
"""
import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnaluwind_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libnaluwind_amd.so is not built (run nalu-wind_amd/build.py)")
lib = C.CDLL(LIB_PATH)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

# status codes
OK, E_ARG, E_STATE, E_HIP, E_NOGRAPHCOL, E_NONFINITE, E_NOCONV, E_ALLOC = (
    0, -1, -2, -3, -4, -5, -6, -7)

# fields
(F_COORDINATES, F_VELOCITY, F_DENSITY, F_PRESSURE, F_DPDX, F_UDIAG,
 F_VISCOSITY, F_DUDX, F_WALL_MASK, F_SCALAR_Q, F_DQDX, F_DIFF_FLUX_COEFF,
 F_DUAL_VOLUME, F_BC_VALUE, E_AREA_VECTOR, E_MASS_FLOW_RATE, E_PECLET_FACTOR,
 E_PECLET_NUMBER, F_PRESSURE_BC, F_UTMP, F_BUOYANCY_SOURCE,
 F_BUOYANCY_SOURCE_MASK) = range(22)
LINSYS_SOLUTION = -1  # nw_solution_update: the linear system's solution

PREC_NONE, PREC_JACOBI, PREC_CHEBYSHEV, PREC_AMG = 0, 1, 2, 3
SOLVE_FP64, SOLVE_FP32_REFINE = 0, 1
X0_ZERO, X0_SOLUTION, X0_LEARNED = 0, 1, 2
MPNN_FP32, MPNN_BF16 = 0, 1
MPNN_LOSS_SOLUTION, MPNN_LOSS_RESIDUAL = 0, 1


class NwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nalu_amd error %d: %s" % (code, msg))
        self.code = code


class Peclet(C.Structure):
    _fields_ = [("form", C.c_int), ("hybrid", C.c_double), ("c1", C.c_double),
                ("c2", C.c_double)]


class ScalarOpts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("alpha_upw", C.c_double),
                ("upw", C.c_double), ("relax", C.c_double),
                ("use_limiter", C.c_int), ("peclet", Peclet),
                ("accumulate", C.c_int)]


class MomentumOpts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("alpha_upw", C.c_double),
                ("upw", C.c_double), ("relax", C.c_double),
                ("include_divU", C.c_double), ("use_limiter", C.c_int),
                ("accumulate", C.c_int), ("extract_diagonal", C.c_int)]


class MdotOpts(C.Structure):
    _fields_ = [("noc", C.c_double), ("interp_together", C.c_double),
                ("balanced_buoyancy", C.c_int), ("gravity", C.c_double * 3)]


class ContinuityOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("gamma1", C.c_double), ("noc", C.c_double),
                ("interp_together", C.c_double),
                ("incompressible", C.c_double), ("accumulate", C.c_int),
                ("balanced_buoyancy", C.c_int), ("gravity", C.c_double * 3)]


class SolverCfg(C.Structure):
    _fields_ = [("kspace", C.c_int), ("max_iters", C.c_int), ("tol", C.c_double),
                ("precond", C.c_int), ("cheb_degree", C.c_int),
                ("cheb_ratio", C.c_double), ("mode", C.c_int),
                ("inner_tol", C.c_double), ("x0", C.c_int),
                ("reuse_preconditioner", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("refinements", C.c_int),
                ("reorthogonalizations", C.c_int), ("converged", C.c_int), ("final_rel", C.c_double),
                ("final_abs", C.c_double), ("bnorm", C.c_double),
                ("ms", C.c_double), ("setup_ms", C.c_double),
                ("precond_levels", C.c_int), ("guess", C.c_int),
                ("x0_rel", C.c_double), ("guess_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MpnnTrainCfg(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("seed", C.c_uint64), ("holdout", C.c_int)]


class ExoInfo(C.Structure):
    _fields_ = [("nnode", C.c_int64), ("nelem", C.c_int64), ("nblock", C.c_int32),
                ("nnodeset", C.c_int32), ("nsideset", C.c_int32),
                ("nnodvar", C.c_int32), ("ntime", C.c_int32), ("format", C.c_int32)]


class ExoSets(C.Structure):
    _fields_ = [("count", C.c_int32), ("ids", C.c_void_p), ("names", C.c_void_p),
                ("offsets", C.c_void_p), ("entries", C.c_void_p),
                ("sides", C.c_void_p)]


EXO_BLOCK, EXO_NODE_SET, EXO_SIDE_SET, EXO_NODAL_VAR = range(4)


def mpnn_train_cfg(**kw):
    o = MpnnTrainCfg(10, 4, 1e-3, 0.9, 0.999, 1e-8, 0, 0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def scalar_defaults(**kw):
    """SolutionOptions defaults (src/SolutionOptions.C:32-52)."""
    o = ScalarOpts(0.0, 1.0, 1.0, 1.0, 0, Peclet(0, 0.0, 2.0, 4.0), 0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def momentum_defaults(**kw):
    o = MomentumOpts(0.0, 1.0, 1.0, 1.0, 0.0, 0, 0, 0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def continuity_defaults(**kw):
    o = ContinuityOpts(1.0, 1.0, 1.0, 1.0, 0.0, 0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def solver_cfg(**kw):
    o = SolverCfg(50, 1000, 1e-8, PREC_JACOBI, 4, 30.0, SOLVE_FP64, 0.0, X0_ZERO, 1)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


_P = C.c_void_p
_sig = {
    "nw_ctx_create": ([C.c_int, C.POINTER(_P)], C.c_int),
    "nw_ctx_destroy": ([_P], C.c_int),
    "nw_ctx_stream": ([_P, C.POINTER(_P)], C.c_int),
    "nw_ctx_synchronize": ([_P], C.c_int),
    "nw_last_error": ([], C.c_char_p),
    "nw_mesh_set": ([_P, C.c_int64, C.c_int64, _i32p], C.c_int),
    "nw_field_upload": ([_P, C.c_int, _f64p, C.c_int64], C.c_int),
    "nw_field_download": ([_P, C.c_int, _f64p, C.c_int64], C.c_int),
    "nw_field_fill": ([_P, C.c_int, C.c_double], C.c_int),
    "nw_field_device_ptr": ([_P, C.c_int, C.POINTER(_P)], C.c_int),
    "nw_mesh_coordinates_changed": ([_P], C.c_int),
    "nw_field_components": ([C.c_int], C.c_int),
    "nw_linsys_build_edge_graph": ([_P, C.c_int], C.c_int),
    "nw_linsys_zero": ([_P], C.c_int),
    "nw_linsys_load_complete": ([_P], C.c_int),
    "nw_linsys_sizes": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "nw_linsys_download": ([_P, _P, _P, _P, _P], C.c_int),
    "nw_linsys_device_arrays": ([_P, C.POINTER(_P), C.POINTER(_P),
                                 C.POINTER(_P), C.POINTER(_P)], C.c_int),
    "nw_linsys_sum_into": ([_P, C.c_int, C.c_int, _i32p, _f64p, _f64p], C.c_int),
    "nw_linsys_reset_rows": ([_P, C.c_int64, _i32p, C.c_int, C.c_int,
                              C.c_double, C.c_double], C.c_int),
    "nw_linsys_apply_dirichlet": ([_P, C.c_int64, _i32p, C.c_int, C.c_int,
                                   C.c_int, C.c_int], C.c_int),
    "nw_fix_pressure_at_node": ([_P, C.c_int32, C.c_double], C.c_int),
    "nw_linsys_set_skipped_rows": ([_P, C.c_int64, _i32p], C.c_int),
    "nw_ctx_set_check_finite": ([_P, C.c_int], C.c_int),
    "nw_linsys_sum_into_device": ([_P, C.c_int, C.c_int, _P, _P, _P], C.c_int),
    "nw_linsys_sum_into_device_noalloc": ([_P, C.c_int, C.c_int, _P, _P, _P],
                                          C.c_int),
    "nw_assemble_scalar": ([_P, C.POINTER(ScalarOpts)], C.c_int),
    "nw_assemble_momentum": ([_P, C.POINTER(MomentumOpts)], C.c_int),
    "nw_assemble_continuity": ([_P, C.POINTER(ContinuityOpts)], C.c_int),
    "nw_momentum_edge_peclet": ([_P, C.POINTER(Peclet)], C.c_int),
    "nw_assemble_momentum_with_peclet": ([_P, C.POINTER(MomentumOpts),
                                          C.POINTER(Peclet)], C.c_int),
    "nw_compute_mdot": ([_P, C.POINTER(MdotOpts)], C.c_int),
    "nw_nodal_grad": ([_P, C.c_int, C.c_int, C.c_int], C.c_int),
    "nw_solution_update": ([_P, C.c_double, C.c_int, C.c_double, C.c_int], C.c_int),
    "nw_project_nodal_velocity": ([_P], C.c_int),
    "nw_mesh_elements": ([_P, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
    "nw_compute_geometry": ([_P], C.c_int),
    "nw_linsys_solve": ([_P, C.POINTER(SolverCfg), C.POINTER(SolveStats)], C.c_int),
    "nw_linsys_solution_upload": ([_P, _f64p], C.c_int),
    "nw_linsys_solution_download": ([_P, _f64p], C.c_int),
    "nw_linsys_solution_device_ptr": ([_P, C.POINTER(_P)], C.c_int),
    "nw_linsys_spmv": ([_P, _P, _P], C.c_int),
    "nw_mpnn_nparams": ([C.c_int, C.c_int], C.c_int64),
    "nw_mpnn_create": ([_P, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)], C.c_int),
    "nw_mpnn_destroy": ([_P], None),
    "nw_mpnn_set_params": ([_P, _f64p, C.c_int64], C.c_int),
    "nw_mpnn_get_params": ([_P, _f64p, C.c_int64], C.c_int),
    "nw_mpnn_record": ([_P], C.c_int),
    "nw_mpnn_samples": ([_P], C.c_int),
    "nw_mpnn_clear_samples": ([_P], C.c_int),
    "nw_mpnn_loss_grad": ([_P, C.c_int, C.POINTER(C.c_double), _P], C.c_int),
    "nw_mpnn_train": ([_P, C.POINTER(MpnnTrainCfg), C.POINTER(C.c_double)], C.c_int),
    "nw_mpnn_predict": ([_P, _f64p, C.POINTER(C.c_double)], C.c_int),
    "nw_mpnn_set_precision": ([_P, C.c_int], C.c_int),
    "nw_mpnn_set_loss": ([_P, C.c_int], C.c_int),
    "nw_linsys_amg_level": ([_P, C.c_int, C.POINTER(C.c_int64),
                             C.POINTER(C.c_int64)], C.c_int),
    "nw_linsys_set_learned_guess": ([_P, _P], C.c_int),
    "nw_mixed_box": ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                      C.POINTER(C.c_int64), C.POINTER(C.c_int64), _P, _P, _P],
                     C.c_int),
    "nw_mixed_edges": ([C.c_int64, _i32p, _i32p, C.c_int64, _P, C.c_int64,
                        C.POINTER(C.c_int64)], C.c_int),
    "nw_mixed_edge_area_vector": ([C.c_int64, _i32p, _i32p, _f64p, _i32p,
                                   C.c_int64, _f64p], C.c_int),
    "nw_fourier_velocity": ([_P, C.c_int, _f64p, _f64p, _f64p, _f64p,
                             C.c_double], C.c_int),
    "nw_hex_box_sizes": ([C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64),
                          C.POINTER(C.c_int64)], C.c_int),
    "nw_hex_box": ([C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                    C.c_uint64, _f64p, _i32p], C.c_int),
    "nw_hex_edges": ([C.c_int64, _i32p, C.c_int64, _i32p, C.c_int64,
                      C.POINTER(C.c_int64), _i32p], C.c_int),
    "nw_hex_edge_area_vector": ([C.c_int64, _i32p, _f64p, _i32p, _i32p,
                                 C.c_int64, _f64p], C.c_int),
    "nw_hex_dual_volume": ([C.c_int64, _i32p, _f64p, C.c_int64, _f64p], C.c_int),
    "nw_nodal_grad_edge": ([C.c_int64, C.c_int64, _i32p, _f64p, _f64p, C.c_int,
                            _f64p, _f64p], C.c_int),
    "nw_mdot_edge": ([C.c_int64, _i32p, _f64p, _f64p, _f64p, _f64p], C.c_int),
    "nw_exo_open": ([C.c_char_p, C.POINTER(_P)], C.c_int),
    "nw_exo_close": ([_P], None),
    "nw_exo_sizes": ([_P, C.POINTER(ExoInfo)], C.c_int),
    "nw_exo_entity": ([_P, C.c_int, C.c_int32, C.POINTER(C.c_int32),
                       C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_char_p,
                       C.c_int32], C.c_int),
    "nw_exo_find": ([_P, C.c_int, C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
    "nw_exo_read_mesh": ([_P, _f64p, _i32p, _i32p, _P], C.c_int),
    "nw_exo_node_set": ([_P, C.c_int32, _i32p], C.c_int),
    "nw_exo_side_set": ([_P, C.c_int32, _i32p, _i32p], C.c_int),
    "nw_exo_read_nodal_var": ([_P, C.c_int32, C.c_int32, _f64p,
                               C.POINTER(C.c_double)], C.c_int),
    "nw_side_nodes": ([C.c_int32, C.c_int32, _i32p], C.c_int),
    "nw_sides_to_nodes": ([C.c_int64, _i32p, _i32p, C.c_int64, C.c_int64, _i32p,
                           _i32p, _P, C.POINTER(C.c_int64)], C.c_int),
    "nw_exo_write": ([C.c_char_p, C.c_char_p, C.c_int64, _f64p, C.c_int64, _i32p,
                      _i32p, C.POINTER(ExoSets), C.POINTER(ExoSets), C.c_int32,
                      _P, _P, C.c_double, _P], C.c_int),
}
for _name, (_a, _r) in _sig.items():
    _f = getattr(lib, _name)
    _f.argtypes = _a
    _f.restype = _r

EXPORTED = sorted(_sig)


def last_error():
    return lib.nw_last_error().decode()


def check(rc):
    if rc != OK:
        raise NwError(rc, last_error())
    return rc


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ mesh

def hex_box(nx, ny, nz, stretch_z=1.0, jitter=0.0, seed=0):
    nn, ne = C.c_int64(), C.c_int64()
    check(lib.nw_hex_box_sizes(nx, ny, nz, C.byref(nn), C.byref(ne)))
    coords = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 8), np.int32)
    check(lib.nw_hex_box(nx, ny, nz, stretch_z, jitter, seed, coords, conn))
    return coords, conn


def mixed_box(n, f_hex=0.6, f_tet=0.3, jitter=0.0, seed=0):
    """Conforming mixed Hex8/Tet4/Wed6/Pyr5 box (transition cells of Pyr5
    and Tet4 about a centre node where Hex8 or Wed6 faces meet cut ones):
    (coords [N][3], type [E], conn [E][8])."""
    nn, ne = C.c_int64(), C.c_int64()
    check(lib.nw_mixed_box(n, f_hex, f_tet, jitter, seed, C.byref(nn),
                           C.byref(ne), None, None, None))
    coords = np.zeros((nn.value, 3))
    types = np.zeros(ne.value, np.int32)
    conn = np.zeros((ne.value, 8), np.int32)
    check(lib.nw_mixed_box(n, f_hex, f_tet, jitter, seed, C.byref(nn),
                           C.byref(ne), coords.ctypes.data_as(_P),
                           types.ctypes.data_as(_P), conn.ctypes.data_as(_P)))
    return coords, types, conn


def mixed_edges(types, conn, nnode):
    types, conn = _i32(types), _i32(conn)
    ne = C.c_int64()
    check(lib.nw_mixed_edges(types.size, types, conn, nnode, None, 0, C.byref(ne)))
    edges = np.zeros((ne.value, 2), np.int32)
    check(lib.nw_mixed_edges(types.size, types, conn, nnode,
                             edges.ctypes.data_as(_P), ne.value, C.byref(ne)))
    return edges


def mixed_edge_area_vector(types, conn, coords, edges):
    out = np.zeros((edges.shape[0], 3))
    check(lib.nw_mixed_edge_area_vector(len(types), _i32(types), _i32(conn),
                                        _f64(coords), _i32(edges),
                                        edges.shape[0], out))
    return out


def hex_edges(conn, nnode):
    conn = _i32(conn)
    nelem = conn.shape[0]
    cap = 3 * nnode + 12 * nelem
    edges = np.zeros((cap, 2), np.int32)
    ee = np.zeros((nelem, 12), np.int32)
    ne = C.c_int64()
    check(lib.nw_hex_edges(nelem, conn, nnode, edges, cap, C.byref(ne), ee))
    return edges[:ne.value].copy(), ee


def hex_edge_area_vector(conn, coords, elem_edges, edges):
    out = np.zeros((edges.shape[0], 3))
    check(lib.nw_hex_edge_area_vector(conn.shape[0], _i32(conn), _f64(coords),
                                      _i32(elem_edges), _i32(edges),
                                      edges.shape[0], out))
    return out


def hex_dual_volume(conn, coords):
    out = np.zeros(coords.shape[0])
    check(lib.nw_hex_dual_volume(conn.shape[0], _i32(conn), _f64(coords),
                                 coords.shape[0], out))
    return out


def nodal_grad_edge(edges, areav, dualvol, phi):
    phi = _f64(phi)
    dim1 = 1 if phi.ndim == 1 else phi.shape[1]
    n = dualvol.shape[0]
    out = np.zeros((n, 3 * dim1))
    check(lib.nw_nodal_grad_edge(n, edges.shape[0], _i32(edges), _f64(areav),
                                 _f64(dualvol), dim1, phi, out))
    return out


def mdot_edge(edges, areav, rho, vel):
    out = np.zeros(edges.shape[0])
    check(lib.nw_mdot_edge(edges.shape[0], _i32(edges), _f64(areav), _f64(rho),
                           _f64(vel), out))
    return out


# ------------------------------------------------------------ Exodus II I/O

def side_nodes(topo, side):
    """Element-local nodes of side `side` of topology `topo`, outward by the
    right-hand rule."""
    out = np.zeros(4, np.int32)
    n = lib.nw_side_nodes(topo, side, out)
    check(min(n, 0))
    return out[:n]


def sides_to_nodes(types, conn, nnode, elem, side):
    """Sorted unique nodes of a list of element sides (a side set)."""
    types, conn = _i32(types), _i32(conn)
    elem, side = _i32(elem), _i32(side)
    cnt = C.c_int64()
    check(lib.nw_sides_to_nodes(types.size, types, conn, nnode, elem.size, elem,
                                side, None, C.byref(cnt)))
    out = np.zeros(cnt.value, np.int32)
    check(lib.nw_sides_to_nodes(types.size, types, conn, nnode, elem.size, elem,
                                side, out.ctypes.data_as(_P), C.byref(cnt)))
    return out


class ExodusMesh:
    """An Exodus II file (netCDF classic or 64-bit-offset) opened for
    reading: header on open, arrays on request."""

    def __init__(self, path):
        self._h = _P()
        check(lib.nw_exo_open(os.fsencode(path), C.byref(self._h)))
        info = ExoInfo()
        check(lib.nw_exo_sizes(self._h, C.byref(info)))
        self.nnode, self.nelem = info.nnode, info.nelem
        self.format, self.ntime = info.format, info.ntime
        self.blocks = [self._entity(EXO_BLOCK, i) for i in range(info.nblock)]
        self.node_sets = [self._entity(EXO_NODE_SET, i) for i in range(info.nnodeset)]
        self.side_sets = [self._entity(EXO_SIDE_SET, i) for i in range(info.nsideset)]
        self.nodal_vars = [self._entity(EXO_NODAL_VAR, i)["name"]
                           for i in range(info.nnodvar)]

    def _entity(self, kind, i):
        eid, cnt, topo = C.c_int32(), C.c_int64(), C.c_int32()
        name = C.create_string_buffer(256)
        check(lib.nw_exo_entity(self._h, kind, i, C.byref(eid), C.byref(cnt),
                                C.byref(topo), name, 256))
        return {"id": eid.value, "count": cnt.value, "topo": topo.value,
                "name": name.value.decode()}

    def _index(self, kind, key):
        if isinstance(key, str):
            idx = C.c_int32()
            check(lib.nw_exo_find(self._h, kind, key.encode(), C.byref(idx)))
            return idx.value
        return int(key)

    def read_mesh(self):
        """(coords [N][3], type [E], conn [E][8], block [E])"""
        coords = np.zeros((self.nnode, 3))
        types = np.zeros(self.nelem, np.int32)
        conn = np.zeros((self.nelem, 8), np.int32)
        block = np.zeros(self.nelem, np.int32)
        check(lib.nw_exo_read_mesh(self._h, coords, types, conn,
                                   block.ctypes.data_as(_P)))
        return coords, types, conn, block

    def node_set(self, key):
        i = self._index(EXO_NODE_SET, key)
        out = np.zeros(self.node_sets[i]["count"], np.int32)
        check(lib.nw_exo_node_set(self._h, i, out))
        return out

    def side_set(self, key):
        """(elem, side) of side set `key` (name or index)"""
        i = self._index(EXO_SIDE_SET, key)
        n = self.side_sets[i]["count"]
        elem, side = np.zeros(n, np.int32), np.zeros(n, np.int32)
        check(lib.nw_exo_side_set(self._h, i, elem, side))
        return elem, side

    def nodal_var(self, key, step=0):
        """(values [N], time) of nodal variable `key` at time step `step`"""
        i = self._index(EXO_NODAL_VAR, key)
        out = np.zeros(self.nnode)
        t = C.c_double()
        check(lib.nw_exo_read_nodal_var(self._h, i, step, out, C.byref(t)))
        return out, t.value

    def close(self):
        if self._h:
            lib.nw_exo_close(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _exo_sets(sets, sides):
    """{name: nodes} or {name: (elem, side)} -> (ExoSets with ids 1.., the
    arrays it points into)"""
    if not sets:
        return None, ()
    names = (C.c_char_p * len(sets))(*[k.encode() for k in sets])
    ents = [_i32(v[0] if sides else v) for v in sets.values()]
    offsets = np.zeros(len(sets) + 1, np.int64)
    offsets[1:] = np.cumsum([e.size for e in ents])
    entries = _i32(np.concatenate(ents))
    sd = _i32(np.concatenate([_i32(v[1]) for v in sets.values()])) if sides else None
    s = ExoSets(len(sets), None, C.cast(names, C.c_void_p).value,
                offsets.ctypes.data, entries.ctypes.data,
                sd.ctypes.data if sides else None)
    return s, (names, offsets, entries, sd)


def exo_write(path, coords, types, conn, node_sets=None, side_sets=None,
              fields=None, time=0.0, title="mesh"):
    """Write a CDF-2 Exodus II file.  node_sets {name: nodes}, side_sets
    {name: (elem, side)}, fields {name: values [N]} at one time step.
    Returns the file's element order (file element j is input element
    order[j]: elements are grouped by topology)."""
    coords, types, conn = _f64(coords), _i32(types), _i32(conn)
    ns, keep_ns = _exo_sets(node_sets, False)
    ss, keep_ss = _exo_sets(side_sets, True)
    fields = fields or {}
    vn = (C.c_char_p * max(len(fields), 1))(*[k.encode() for k in fields])
    vals = (_f64(np.stack([np.asarray(v, np.float64) for v in fields.values()]))
            if fields else None)
    order = np.zeros(types.size, np.int64)
    check(lib.nw_exo_write(os.fsencode(path), title.encode(), coords.shape[0],
                           coords, types.size, types, conn,
                           C.byref(ns) if ns is not None else None,
                           C.byref(ss) if ss is not None else None, len(fields),
                           C.cast(vn, C.c_void_p) if fields else None,
                           vals.ctypes.data_as(_P) if fields else None, time,
                           order.ctypes.data_as(_P)))
    del keep_ns, keep_ss
    return order


# ------------------------------------------------------------------ context

_hip = None


def _hip_rt():
    """The HIP runtime the library links (device buffers for the calls that
    take device arrays)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(_P), C.c_size_t]
        _hip.hipFree.argtypes = [_P]
        _hip.hipMemcpy.argtypes = [_P, _P, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceBuffer:
    """A device copy of a numpy array (hipMalloc + hipMemcpy)."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr)
        self.ptr = _P()
        rt = _hip_rt()
        nb = max(self.arr.nbytes, 1)
        if rt.hipMalloc(C.byref(self.ptr), nb) != 0:
            raise NwError(E_ALLOC, "hipMalloc of %d bytes failed" % nb)
        self.ptr = self.ptr.value
        if self.arr.nbytes:
            assert rt.hipMemcpy(self.ptr, self.arr.ctypes.data, self.arr.nbytes, 1) == 0

    def download(self):
        out = np.empty_like(self.arr)
        _hip_rt().hipDeviceSynchronize()
        if out.nbytes:
            assert _hip_rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            _hip_rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU context: mesh, fields, one linear system, solver workspace."""

    def __init__(self, device=0):
        self.h = _P()
        check(lib.nw_ctx_create(device, C.byref(self.h)))
        self.nnode = 0
        self.nedge = 0
        self._nets = weakref.WeakSet()

    def close(self):
        if self.h:
            for net in list(self._nets):  # networks go before their context
                net.close()
            lib.nw_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stream(self):
        s = _P()
        check(lib.nw_ctx_stream(self.h, C.byref(s)))
        return s.value

    def synchronize(self):
        check(lib.nw_ctx_synchronize(self.h))

    def mesh_set(self, nnode, edges):
        edges = _i32(edges).reshape(-1, 2)
        check(lib.nw_mesh_set(self.h, nnode, edges.shape[0], edges))
        self.nnode, self.nedge = nnode, edges.shape[0]

    def field_upload(self, fid, arr):
        a = _f64(arr).ravel()
        check(lib.nw_field_upload(self.h, fid, a, a.size))

    def field_download(self, fid):
        nc = lib.nw_field_components(fid)
        edge = fid in (E_AREA_VECTOR, E_MASS_FLOW_RATE, E_PECLET_FACTOR,
                       E_PECLET_NUMBER)
        n = (self.nedge if edge else self.nnode) * nc
        out = np.zeros(n)
        check(lib.nw_field_download(self.h, fid, out, n))
        return out.reshape(-1, nc) if nc > 1 else out

    def field_fill(self, fid, value):
        check(lib.nw_field_fill(self.h, fid, value))

    def field_ptr(self, fid):
        p = _P()
        check(lib.nw_field_device_ptr(self.h, fid, C.byref(p)))
        return p.value

    def build_graph(self, ndof):
        check(lib.nw_linsys_build_edge_graph(self.h, ndof))
        self.ndof = ndof

    def zero(self):
        check(lib.nw_linsys_zero(self.h))

    def load_complete(self):
        check(lib.nw_linsys_load_complete(self.h))

    def sizes(self):
        a, b = C.c_int64(), C.c_int64()
        check(lib.nw_linsys_sizes(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def download(self):
        """(row_ptr, cols, vals, rhs) in the Tpetra point-CSR layout."""
        nrows, nnz = self.sizes()
        rp = np.zeros(nrows + 1, np.int64)
        cols = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz)
        rhs = np.zeros(nrows)
        check(lib.nw_linsys_download(self.h, rp.ctypes.data, cols.ctypes.data,
                                     vals.ctypes.data, rhs.ctypes.data))
        return rp, cols, vals, rhs

    def device_arrays(self):
        ps = [_P() for _ in range(4)]
        check(lib.nw_linsys_device_arrays(self.h, *[C.byref(p) for p in ps]))
        return [p.value for p in ps]

    def sum_into(self, nodes, lhs, rhs):
        nodes = _i32(nodes)
        nb = nodes.shape[0] if nodes.ndim == 2 else 1
        nent = nodes.shape[-1]
        check(lib.nw_linsys_sum_into(self.h, nb, nent, nodes.ravel(),
                                     _f64(lhs).ravel(), _f64(rhs).ravel()))

    def set_check_finite(self, on=True):
        """Assemblies return E_NONFINITE when the system holds NaN/Inf."""
        check(lib.nw_ctx_set_check_finite(self.h, int(on)))

    def sum_into_device(self, d_nodes, d_lhs, d_rhs, nblocks, nent):
        """sumInto of device arrays (DeviceBuffer or raw pointers)."""
        ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
        check(lib.nw_linsys_sum_into_device(self.h, nblocks, nent, ptr(d_nodes),
                                            ptr(d_lhs), ptr(d_rhs)))

    def sum_into_device_noalloc(self, d_nodes, d_lhs, d_rhs, nblocks, nent):
        """sum_into_device with the library's fixed scratch: allocates
        nothing (for callers that cannot allocate device memory)."""
        ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else b
        check(lib.nw_linsys_sum_into_device_noalloc(self.h, nblocks, nent,
                                                    ptr(d_nodes), ptr(d_lhs),
                                                    ptr(d_rhs)))

    def reset_rows(self, nodes, begin, end, diag=0.0, rhs_residual=0.0):
        nodes = _i32(nodes)
        check(lib.nw_linsys_reset_rows(self.h, nodes.size, nodes, begin, end,
                                       diag, rhs_residual))

    def apply_dirichlet(self, nodes, sol_field, bc_field, begin, end):
        nodes = _i32(nodes)
        check(lib.nw_linsys_apply_dirichlet(self.h, nodes.size, nodes, sol_field,
                                            bc_field, begin, end))

    def set_skipped_rows(self, nodes):
        """Skipped rows of the hypre backend (buildDirichletNodeGraph)."""
        nodes = _i32(nodes)
        check(lib.nw_linsys_set_skipped_rows(self.h, nodes.size, nodes))

    def fix_pressure_at_node(self, node, pref):
        check(lib.nw_fix_pressure_at_node(self.h, int(node), pref))

    def assemble_scalar(self, opts):
        check(lib.nw_assemble_scalar(self.h, C.byref(opts)))

    def assemble_momentum(self, opts):
        check(lib.nw_assemble_momentum(self.h, C.byref(opts)))

    def assemble_continuity(self, opts):
        check(lib.nw_assemble_continuity(self.h, C.byref(opts)))

    def momentum_edge_peclet(self, peclet):
        check(lib.nw_momentum_edge_peclet(self.h, C.byref(peclet)))

    def assemble_momentum_with_peclet(self, opts, peclet):
        """MomentumEdgePecletAlg + MomentumEdgeSolverAlg in one pass."""
        check(lib.nw_assemble_momentum_with_peclet(self.h, C.byref(opts),
                                                   C.byref(peclet)))

    def compute_mdot(self, noc=1.0, interp_together=1.0, gravity=None):
        """MdotEdgeAlg: NW_E_MASS_FLOW_RATE from the node fields; with
        `gravity` (3 values) the balanced buoyancy force is on and the
        F_BUOYANCY_SOURCE / F_BUOYANCY_SOURCE_MASK fields are read."""
        o = MdotOpts(noc, interp_together, 0)
        if gravity is not None:
            o.balanced_buoyancy = 1
            o.gravity = (C.c_double * 3)(*[float(g) for g in gravity])
        check(lib.nw_compute_mdot(self.h, C.byref(o)))

    def nodal_grad(self, field_in, field_out, accumulate=0):
        """NodalGradEdgeAlg: field_out = (or +=, with accumulate) the
        edge-based gradient of field_in."""
        check(lib.nw_nodal_grad(self.h, field_in, field_out, int(accumulate)))

    def solution_update(self, delta_frac, delta, field_frac, field):
        """EquationSystem::solution_update: field = delta_frac * delta +
        field_frac * field (delta = LINSYS_SOLUTION: the solved increment)."""
        check(lib.nw_solution_update(self.h, delta_frac, delta, field_frac, field))

    def project_nodal_velocity(self):
        """LowMachEquationSystem::project_nodal_velocity."""
        check(lib.nw_project_nodal_velocity(self.h))

    def mesh_elements(self, types, conn):
        """The mesh's elements for the device geometry: types [nelem]
        (0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5; None = all Hex8), conn [nelem][8]."""
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 8)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        check(lib.nw_mesh_elements(
            self.h, conn.shape[0],
            None if t is None else t.ctypes.data_as(C.c_void_p),
            conn.ctypes.data_as(C.c_void_p)))

    def compute_geometry(self):
        """GeometryInteriorAlg on the device: NW_E_AREA_VECTOR and
        NW_F_DUAL_VOLUME from the coordinates."""
        check(lib.nw_compute_geometry(self.h))

    def amg_levels(self):
        """[(rows, nnz)] of the AMG hierarchy built by the last AMG solve."""
        out = []
        r, z = C.c_int64(), C.c_int64()
        while lib.nw_linsys_amg_level(self.h, len(out), C.byref(r), C.byref(z)) == OK:
            out.append((r.value, z.value))
        return out

    def solve(self, cfg, raise_on_noconv=True):
        st = SolveStats()
        rc = lib.nw_linsys_solve(self.h, C.byref(cfg), C.byref(st))
        if rc != OK and not (rc == E_NOCONV and not raise_on_noconv):
            raise NwError(rc, last_error())
        return st

    def solution(self):
        nrows, _ = self.sizes()
        out = np.zeros(nrows)
        check(lib.nw_linsys_solution_download(self.h, out))
        return out

    def set_solution(self, x):
        check(lib.nw_linsys_solution_upload(self.h, _f64(x)))

    def solution_ptr(self):
        p = _P()
        check(lib.nw_linsys_solution_device_ptr(self.h, C.byref(p)))
        return p.value

    def spmv_device(self, d_x, d_y):
        check(lib.nw_linsys_spmv(self.h, d_x, d_y))

    def fourier_velocity(self, modes, t):
        """Device-generated velocity field sum_q b_q sin(k_q.x + ph_q + w_q t)
        from modes = (k [m][3], b [m][3], phase [m], omega [m])."""
        K, B, PH, OM = (_f64(a).ravel() for a in modes)
        check(lib.nw_fourier_velocity(self.h, PH.size, K, B, PH, OM, t))

    def set_learned_guess(self, net):
        check(lib.nw_linsys_set_learned_guess(self.h, net.h if net else None))


class Mpnn:
    """Learned initial guess of the pressure solve on a context's system."""

    def __init__(self, ctx, hidden=64, layers=3, seed=0):
        self.ctx = ctx
        self.h = _P()
        check(lib.nw_mpnn_create(ctx.h, hidden, layers, seed, C.byref(self.h)))
        ctx._nets.add(self)
        self.hidden, self.layers = hidden, layers
        self.nparams = lib.nw_mpnn_nparams(hidden, layers)
        self._precision = MPNN_BF16
        self._loss = MPNN_LOSS_RESIDUAL

    @property
    def loss(self):
        """Training loss: MPNN_LOSS_RESIDUAL (default) or MPNN_LOSS_SOLUTION."""
        return self._loss

    @loss.setter
    def loss(self, kind):
        check(lib.nw_mpnn_set_loss(self.h, kind))
        self._loss = kind

    @property
    def precision(self):
        """Arithmetic of the prediction: MPNN_BF16 (default) or MPNN_FP32."""
        return self._precision

    @precision.setter
    def precision(self, p):
        check(lib.nw_mpnn_set_precision(self.h, p))
        self._precision = p

    def close(self):
        if self.h:
            lib.nw_mpnn_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.nparams)
        check(lib.nw_mpnn_get_params(self.h, p, self.nparams))
        return p

    @params.setter
    def params(self, p):
        p = _f64(p)
        check(lib.nw_mpnn_set_params(self.h, p, p.size))

    def record(self):
        check(lib.nw_mpnn_record(self.h))

    @property
    def samples(self):
        return lib.nw_mpnn_samples(self.h)

    def clear_samples(self):
        check(lib.nw_mpnn_clear_samples(self.h))

    def loss_grad(self, sample):
        loss = C.c_double()
        g = np.zeros(self.nparams)
        check(lib.nw_mpnn_loss_grad(self.h, sample, C.byref(loss),
                                    g.ctypes.data_as(_P)))
        return loss.value, g

    def train(self, cfg):
        loss = C.c_double()
        check(lib.nw_mpnn_train(self.h, C.byref(cfg), C.byref(loss)))
        return loss.value

    def predict(self):
        nrows, _ = self.ctx.sizes()
        x = np.zeros(nrows)
        sc = C.c_double()
        check(lib.nw_mpnn_predict(self.h, x, C.byref(sc)))
        return x, sc.value

"""feelpp_amd: MI355X-native form assembly and Jacobi-CG solve for Lagrange Pk
on simplices, Python bindings over the C-ABI of libfeelpp_hip.so.

The native library must be built (python -c "import __graft_entry__ as g; g.build()");
nothing here falls back to a CPU implementation: if the library or a GPU is
missing, the calls raise.
"""
import atexit
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfeelpp_hip.so")
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

FP64, FP32, FP64_FP32 = 1, 2, 3
MASS, DIFFUSION, ELASTICITY = 1, 2, 4
MATRICES_AUTO, MATRICES_FUSED, MATRICES_MFMA = 0, 1, 2
SRC_CONST, SRC_SINSIN, SRC_VECTOR = 0, 1, 2
ELIMINATION, ELIMINATION_SYMMETRIC = 0, 1
INNER_AUTO, INNER_ASSEMBLED, INNER_ELEMENT = 0, 1, 2
SAME_NONZERO_PATTERN, DIFFERENT_NONZERO_PATTERN, SUBSET_NONZERO_PATTERN = 0, 1, 3
NORM_1, NORM_FROBENIUS, NORM_INF = 0, 2, 3
VEC_SUM, VEC_L1, VEC_L2, VEC_MIN, VEC_MAX, VEC_LINF = 0, 1, 2, 3, 4, 5


class FeelHipError(RuntimeError):
    pass


class BilinearDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("quad_order", C.c_int), ("precision", C.c_int),
                ("diff_scale", C.c_double), ("diff_coef", _vp), ("mass_scale", C.c_double),
                ("mass_coef", _vp), ("young", _vp), ("poisson", _vp), ("matrices", C.c_int)]


class LinearDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("quad_order", C.c_int), ("params", C.c_double * 4)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dtol", C.c_double), ("maxit", C.c_int),
                ("precision", C.c_int), ("inner_rtol", C.c_double), ("inner_maxit", C.c_int),
                ("max_outer", C.c_int), ("check_every", C.c_int), ("inner_operator", C.c_int)]


class SolveData(C.Structure):
    _fields_ = [("converged", C.c_int), ("its", C.c_int), ("residual", C.c_double), ("reason", C.c_int),
                ("outer_its", C.c_int), ("rnorm0", C.c_double), ("fp64_its", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("inner_rtol", C.c_double), ("maxit", C.c_int),
                ("inner_maxit", C.c_int), ("max_outer", C.c_int), ("precision", C.c_int), ("hbm_budget", C.c_int64),
                ("use_graph", C.c_int), ("group_spmv", C.c_int), ("backfill", C.c_int)]


# exported symbols of the two C-ABIs: name -> (restype, argtypes)
_SIGS = {
    # host
    "feel_host_last_error": (C.c_char_p, []),
    "feel_mesh_unit_square": (C.c_int, [C.c_int, C.c_double, C.c_uint64, _f64p, _i32p]),
    "feel_mesh_kuhn_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_uint64, _f64p, _i32p]),
    "feel_mesh_write_msh22": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _f64p, C.c_int64, _i32p]),
    "feel_mesh_read_msh22": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int64), _vp, C.POINTER(C.c_int64), _vp]),
    "feel_mesh_write_msh22_solution": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _f64p, C.c_int64, _i32p, C.c_int,
                                                 C.c_int, _i32p, _f64p, C.c_char_p, C.c_double, C.c_int]),
    "feel_field_lognormal": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_uint64, _f64p]),
    "feel_field_uniform": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_int, C.c_uint64, _f64p]),
    "feel_doftable_build": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, _i32p, _i32p]),
    "feel_graph_build": (C.c_int, [C.c_int, C.c_int64, C.c_int64, _i32p, C.POINTER(_vp)]),
    "feel_graph_nnz": (C.c_int64, [_vp]),
    "feel_graph_copy": (C.c_int, [_vp, _i32p, _i32p]),
    "feel_graph_free": (None, [_vp]),
    "feel_boundary_dofs": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int64, _f64p, C.c_int64, _i32p, _i32p,
                                       C.c_int64, C.c_int, _u8p]),
    "feel_lagrange_nodes": (C.c_int, [C.c_int, C.c_int, _vp]),
    "feel_quadrature": (C.c_int, [C.c_int, C.c_int, _vp, _vp]),
    "feel_reference_tensors": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "feel_cg_condition_estimate": (C.c_int, [_vp, C.c_int, _vp]),
    "feel_cg_condition_exceeds": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    # device
    "feel_hip_last_error": (C.c_char_p, []),
    "feel_hip_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "feel_hip_ctx_destroy": (C.c_int, [_vp]),
    "feel_hip_ctx_synchronize": (C.c_int, [_vp]),
    "feel_hip_ctx_stream": (_vp, [_vp]),
    "feel_hip_ctx_device_bytes": (C.c_int64, [_vp]),
    "feel_hip_ctx_set_timing": (C.c_int, [_vp, C.c_int]),
    "feel_hip_ctx_timing": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "feel_hip_ctx_reset_timing": (C.c_int, [_vp]),
    "feel_hip_space_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int64, _f64p, C.c_int64, _i32p, _i32p,
                                        C.c_int64, C.POINTER(_vp)]),
    "feel_hip_space_create_from_mesh": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int64, _f64p, C.c_int64,
                                                  _i32p, C.POINTER(_vp)]),
    "feel_hip_space_elem2dof": (C.c_int, [_vp, _i32p]),
    "feel_hip_space_destroy": (C.c_int, [_vp]),
    "feel_hip_space_ndof": (C.c_int64, [_vp]),
    "feel_hip_space_nelements": (C.c_int64, [_vp]),
    "feel_hip_graph_create": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, C.POINTER(_vp)]),
    "feel_hip_graph_create_from_space": (C.c_int, [_vp, C.POINTER(_vp)]),
    "feel_hip_graph_pattern": (C.c_int, [_vp, _i32p, _i32p]),
    "feel_hip_graph_destroy": (C.c_int, [_vp]),
    "feel_hip_graph_nnz": (C.c_int64, [_vp]),
    "feel_hip_mat_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "feel_hip_mat_destroy": (C.c_int, [_vp]),
    "feel_hip_mat_zero": (C.c_int, [_vp]),
    "feel_hip_mat_add_matrix": (C.c_int, [_vp, C.c_int, _i32p, C.c_int, _i32p, _f64p]),
    "feel_hip_mat_get_values": (C.c_int, [_vp, C.c_int, _vp]),
    "feel_hip_mat_set_values": (C.c_int, [_vp, _f64p]),
    "feel_hip_mat_mult": (C.c_int, [_vp, _vp, _vp]),
    "feel_hip_mat_scale": (C.c_int, [_vp, C.c_double]),
    "feel_hip_mat_axpy": (C.c_int, [_vp, C.c_double, _vp, C.c_int]),
    "feel_hip_mat_copy": (C.c_int, [_vp, _vp]),
    "feel_hip_mat_diagonal": (C.c_int, [_vp, _vp]),
    "feel_hip_mat_transpose": (C.c_int, [_vp, _vp]),
    "feel_hip_mat_norm": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    "feel_hip_mat_energy": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_double)]),
    "feel_hip_mat_set_matrix": (C.c_int, [_vp, C.c_int, _i32p, C.c_int, _i32p, _f64p]),
    "feel_hip_mat_get_entries": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f64p]),
    "feel_hip_mat_add_matrix_noalloc": (C.c_int, [_vp, C.c_int, _i32p, C.c_int, _i32p, _f64p]),
    "feel_hip_vec_create": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp)]),
    "feel_hip_vec_destroy": (C.c_int, [_vp]),
    "feel_hip_vec_set": (C.c_int, [_vp, _f64p]),
    "feel_hip_vec_get": (C.c_int, [_vp, _f64p]),
    "feel_hip_vec_zero": (C.c_int, [_vp]),
    "feel_hip_vec_dot": (C.c_int, [_vp, _vp, C.POINTER(C.c_double)]),
    "feel_hip_vec_axpy": (C.c_int, [_vp, C.c_double, _vp]),
    "feel_hip_vec_scale": (C.c_int, [_vp, C.c_double]),
    "feel_hip_vec_copy": (C.c_int, [_vp, _vp]),
    "feel_hip_vec_set_rows": (C.c_int, [_vp, _vp, _vp]),
    "feel_hip_vec_size": (C.c_int64, [_vp]),
    "feel_hip_vec_reduce": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    "feel_hip_vec_set_constant": (C.c_int, [_vp, C.c_double]),
    "feel_hip_vec_shift": (C.c_int, [_vp, C.c_double]),
    "feel_hip_vec_set_indexed": (C.c_int, [_vp, C.c_int64, _i32p, _f64p]),
    "feel_hip_vec_add_indexed": (C.c_int, [_vp, C.c_int64, _i32p, _f64p]),
    "feel_hip_vec_mult_add": (C.c_int, [_vp, _vp, _vp]),
    "feel_hip_vec_device_ptr": (_vp, [_vp]),
    "feel_hip_assemble_bilinear": (C.c_int, [_vp, C.POINTER(BilinearDesc), C.c_int]),
    "feel_hip_assemble_linear": (C.c_int, [_vp, C.POINTER(LinearDesc), _vp, C.c_int]),
    "feel_hip_rows_create": (C.c_int, [_vp, C.c_int64, C.c_int64, _i32p, C.POINTER(_vp)]),
    "feel_hip_rows_destroy": (C.c_int, [_vp]),
    "feel_hip_mat_zero_rows": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double]),
    "feel_hip_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "feel_hip_solve_pcg": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveData)]),
    "feel_hip_graph_spmv_shapes": (C.c_int, []),
    "feel_hip_graph_set_spmv_shape": (C.c_int, [_vp, C.c_int, C.c_int]),
    "feel_hip_graph_spmv_shape": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int)]),
    "feel_hip_norm_l2": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_double)]),
    "feel_hip_norm_semi_h1": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_double)]),
    "feel_hip_batch_opts_default": (None, [C.POINTER(BatchOpts)]),
    "feel_hip_batch_create": (C.c_int, [_vp, C.POINTER(BatchOpts), C.POINTER(_vp)]),
    "feel_hip_batch_destroy": (C.c_int, [_vp]),
    "feel_hip_batch_add_family": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), _vp, _vp, C.POINTER(C.c_int)]),
    "feel_hip_batch_submit": (C.c_int, [_vp, C.c_int, _f64p, C.POINTER(C.c_int64)]),
    "feel_hip_batch_step": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "feel_hip_batch_poll": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(SolveData)]),
    "feel_hip_batch_solution": (C.c_int, [_vp, C.c_int64, _vp]),
    "feel_hip_batch_device_bytes": (C.c_int64, [_vp]),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise FeelHipError(f"{LIBPATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, host=False):
    if rc < 0:
        msg = (lib().feel_host_last_error() if host else lib().feel_hip_last_error()) or b""
        raise FeelHipError(f"status {rc}: {msg.decode(errors='replace')}")
    return rc


# ============================================================ host side
def mesh_unit_square(n, jitter=0.0, seed=0):
    coords = np.zeros(((n + 1) ** 2, 2))
    elems = np.zeros((2 * n * n, 3), np.int32)
    _check(lib().feel_mesh_unit_square(n, jitter, seed, coords, elems), host=True)
    return coords, elems


def mesh_kuhn_box(nx, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    elems = np.zeros((6 * nx * ny * nz, 4), np.int32)
    _check(lib().feel_mesh_kuhn_box(nx, ny, nz, lx, ly, lz, jitter, seed, coords, elems), host=True)
    return coords, elems


def write_msh22(path, coords, elems):
    dim = coords.shape[1]
    _check(lib().feel_mesh_write_msh22(path.encode(), dim, coords.shape[0], np.ascontiguousarray(coords),
                                       elems.shape[0], np.ascontiguousarray(elems, dtype=np.int32)), host=True)


def write_msh22_solution(path, coords, elems, order, elem2dof, values, ncomp=1, name="u", time=0.0, step=0):
    """Mesh + $ElementNodeData of a Pk field (vertex values), as Feel++'s Gmsh exporter writes it."""
    dim = coords.shape[1]
    _check(lib().feel_mesh_write_msh22_solution(
        path.encode(), dim, coords.shape[0], np.ascontiguousarray(coords), elems.shape[0],
        np.ascontiguousarray(elems, dtype=np.int32), order, ncomp, np.ascontiguousarray(elem2dof, dtype=np.int32),
        np.ascontiguousarray(values, dtype=np.float64), name.encode(), time, step), host=True)


def read_msh22(path, dim):
    nv, ne = C.c_int64(), C.c_int64()
    _check(lib().feel_mesh_read_msh22(path.encode(), dim, C.byref(nv), None, C.byref(ne), None), host=True)
    coords = np.zeros((nv.value, dim))
    elems = np.zeros((ne.value, dim + 1), np.int32)
    _check(lib().feel_mesh_read_msh22(path.encode(), dim, C.byref(nv), coords.ctypes.data_as(_vp), C.byref(ne),
                                      elems.ctypes.data_as(_vp)), host=True)
    return coords, elems


def lognormal(n, mu, sigma, seed):
    out = np.zeros(n)
    _check(lib().feel_field_lognormal(n, mu, sigma, seed, out), host=True)
    return out


def uniform(n, lo, hi, seed, log=False):
    out = np.zeros(n)
    _check(lib().feel_field_uniform(n, lo, hi, int(log), seed, out), host=True)
    return out


def nloc(dim, order):
    return (order + 1) * (order + 2) // 2 if dim == 2 else (order + 1) * (order + 2) * (order + 3) // 6


def doftable(dim, order, ncomp, elems, nv):
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    e2d = np.zeros((elems.shape[0], nloc(dim, order) * ncomp), np.int32)
    ndof = _check(lib().feel_doftable_build(dim, order, ncomp, nv, elems.shape[0], elems, e2d), host=True)
    return e2d, int(ndof)


def graph(elem2dof, ndof):
    e2d = np.ascontiguousarray(elem2dof, dtype=np.int32)
    h = _vp()
    _check(lib().feel_graph_build(e2d.shape[1], e2d.shape[0], ndof, e2d, C.byref(h)), host=True)
    try:
        nnz = lib().feel_graph_nnz(h)
        ia = np.zeros(ndof + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        _check(lib().feel_graph_copy(h, ia, ja), host=True)
    finally:
        lib().feel_graph_free(h)
    return ia, ja


def boundary_dofs(dim, order, ncomp, coords, elems, elem2dof, ndof, sel=0):
    flags = np.zeros(ndof, np.uint8)
    _check(lib().feel_boundary_dofs(dim, order, ncomp, coords.shape[0], np.ascontiguousarray(coords),
                                    elems.shape[0], np.ascontiguousarray(elems, dtype=np.int32),
                                    np.ascontiguousarray(elem2dof, dtype=np.int32), ndof, sel, flags), host=True)
    return flags


def lagrange_nodes(dim, order):
    n = _check(lib().feel_lagrange_nodes(dim, order, None), host=True)
    out = np.zeros((n, dim))
    lib().feel_lagrange_nodes(dim, order, out.ctypes.data_as(_vp))
    return out


def quadrature(dim, order):
    n = _check(lib().feel_quadrature(dim, order, None, None), host=True)
    pts, w = np.zeros((n, dim)), np.zeros(n)
    lib().feel_quadrature(dim, order, pts.ctypes.data_as(_vp), w.ctypes.data_as(_vp))
    return pts, w


def reference_tensors(dim, order, qorder):
    n = nloc(dim, order)
    R, M, m = np.zeros((dim * dim, n, n)), np.zeros((n, n)), np.zeros(n)
    _check(lib().feel_reference_tensors(dim, order, qorder, R.ctypes.data_as(_vp), M.ctypes.data_as(_vp),
                                        m.ctypes.data_as(_vp)), host=True)
    return R, M, m


def cg_condition_estimate(gamma, pap):
    """Condition number of D^{-1} A from the coefficients of a CG run:
    gamma_j = (r_j, z_j) for j = 0..k and pAp_j = (p_j, A p_j) for j < k."""
    gamma, pap = np.asarray(gamma, dtype=np.float64), np.asarray(pap, dtype=np.float64)
    k = len(pap)
    h = np.ones(2 * (k + 1))
    h[0::2][:len(gamma)] = gamma[:k + 1]
    h[1:2 * k:2] = pap
    out = C.c_double()
    _check(lib().feel_cg_condition_estimate(h.ctypes.data_as(_vp), k, C.byref(out)), host=True)
    return out.value


def cg_condition_exceeds(gamma, pap, limit):
    """Whether cg_condition_estimate(gamma, pap) exceeds limit, decided from
    the largest eigenvalue to 1 % and one Sturm count (the FP32 solve's test)."""
    gamma, pap = np.asarray(gamma, dtype=np.float64), np.asarray(pap, dtype=np.float64)
    k = len(pap)
    h = np.ones(2 * (k + 1))
    h[0::2][:len(gamma)] = gamma[:k + 1]
    h[1:2 * k:2] = pap
    out = C.c_int()
    _check(lib().feel_cg_condition_exceeds(h.ctypes.data_as(_vp), k, float(limit), C.byref(out)), host=True)
    return bool(out.value)


# ============================================================ device side
_exiting = False


def _mark_exiting():
    global _exiting
    _exiting = True


# At interpreter exit, objects caught in reference cycles are finalised in no
# particular order; the process is ending, so handles are left to it.
atexit.register(_mark_exiting)


class _Handle:
    _destroy = None

    def __init__(self, h, keep=()):
        self.h = h
        self._keep = keep

    def close(self):
        if self.h:
            getattr(lib(), self._destroy)(self.h)
            self.h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass


class Context(_Handle):
    _destroy = "feel_hip_ctx_destroy"

    def __init__(self, device=0):
        h = _vp()
        _check(lib().feel_hip_ctx_create(device, C.byref(h)))
        super().__init__(h)

    def synchronize(self):
        _check(lib().feel_hip_ctx_synchronize(self.h))

    @property
    def stream(self):
        return lib().feel_hip_ctx_stream(self.h)

    def device_bytes(self):
        return lib().feel_hip_ctx_device_bytes(self.h)

    def set_timing(self, period):
        """0: off; N: time every N-th launch of each timed kernel (True = every launch)"""
        _check(lib().feel_hip_ctx_set_timing(self.h, int(period)))

    def reset_timing(self):
        _check(lib().feel_hip_ctx_reset_timing(self.h))

    def timing(self, name):
        n, ms = C.c_int64(), C.c_double()
        _check(lib().feel_hip_ctx_timing(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def vector(self, n, values=None):
        v = Vector(self, n)
        if values is not None:
            v.set(values)
        return v


class Vector(_Handle):
    _destroy = "feel_hip_vec_destroy"

    def __init__(self, ctx, n):
        h = _vp()
        _check(lib().feel_hip_vec_create(ctx.h, n, C.byref(h)))
        super().__init__(h, keep=(ctx,))
        self.n = n

    def set(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.shape == (self.n,)
        _check(lib().feel_hip_vec_set(self.h, a))

    def get(self):
        out = np.zeros(self.n)
        _check(lib().feel_hip_vec_get(self.h, out))
        return out

    def zero(self):
        _check(lib().feel_hip_vec_zero(self.h))

    def dot(self, other):
        out = C.c_double()
        _check(lib().feel_hip_vec_dot(self.h, other.h, C.byref(out)))
        return out.value

    def axpy(self, a, x):
        """self += a x (Vector::add(a, v))"""
        _check(lib().feel_hip_vec_axpy(self.h, a, x.h))

    def scale(self, a):
        _check(lib().feel_hip_vec_scale(self.h, a))

    def copy_from(self, src):
        _check(lib().feel_hip_vec_copy(self.h, src.h))

    def set_rows(self, rows, values=None):
        """self[r] = values[r] (0 without values) for the rows of a Rows set"""
        _check(lib().feel_hip_vec_set_rows(self.h, rows.h, values.h if values else None))

    def reduce(self, kind):
        out = C.c_double()
        _check(lib().feel_hip_vec_reduce(self.h, kind, C.byref(out)))
        return out.value

    def sum(self):
        return self.reduce(VEC_SUM)

    def min(self):
        return self.reduce(VEC_MIN)

    def max(self):
        return self.reduce(VEC_MAX)

    def l1_norm(self):
        return self.reduce(VEC_L1)

    def l2_norm(self):
        return self.reduce(VEC_L2)

    def linfty_norm(self):
        return self.reduce(VEC_LINF)

    def set_constant(self, s):
        _check(lib().feel_hip_vec_set_constant(self.h, s))

    def shift(self, s):
        """self += s (Vector::add(s))"""
        _check(lib().feel_hip_vec_shift(self.h, s))

    def set_indexed(self, idx, vals):
        """Vector::setVector(i, n, v): self[idx[k]] = vals[k]"""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        assert idx.shape == vals.shape
        _check(lib().feel_hip_vec_set_indexed(self.h, idx.shape[0], idx, vals))

    def add_indexed(self, idx, vals):
        """Vector::addVector(i, n, v): self[idx[k]] += vals[k]"""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        assert idx.shape == vals.shape
        _check(lib().feel_hip_vec_add_indexed(self.h, idx.shape[0], idx, vals))

    def mult_add(self, A, x):
        """Vector::addVector(V, A): self += A x"""
        _check(lib().feel_hip_vec_mult_add(self.h, A.h, x.h))


class Space(_Handle):
    """A mesh plus the DofTable of Pch<order> (ncomp=1) or Pchv<order> (ncomp=dim), on the device."""
    _destroy = "feel_hip_space_destroy"

    def __init__(self, ctx, coords, elems, order, ncomp=1, elem2dof=None, ndof=None):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        dim = coords.shape[1]
        if elem2dof is None:
            elem2dof, ndof = doftable(dim, order, ncomp, elems, coords.shape[0])
        elem2dof = np.ascontiguousarray(elem2dof, dtype=np.int32)
        h = _vp()
        _check(lib().feel_hip_space_create(ctx.h, dim, order, ncomp, coords.shape[0], coords, elems.shape[0], elems,
                                           elem2dof, ndof, C.byref(h)))
        super().__init__(h, keep=(ctx,))
        self.ctx, self.dim, self.order, self.ncomp = ctx, dim, order, ncomp
        self.coords, self.elems, self.elem2dof, self.ndof = coords, elems, elem2dof, ndof
        self.ne = elems.shape[0]

    @classmethod
    def from_mesh(cls, ctx, coords, elems, order, ncomp=1):
        """The DofTable built on the device from the mesh (Feel++'s first-touch
        numbering); elem2dof is copied back for host-side helpers."""
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        dim = coords.shape[1]
        h = _vp()
        _check(lib().feel_hip_space_create_from_mesh(ctx.h, dim, order, ncomp, coords.shape[0], coords,
                                                     elems.shape[0], elems, C.byref(h)))
        self = cls.__new__(cls)
        _Handle.__init__(self, h, keep=(ctx,))
        nloc = (order + 1) * (order + 2) // 2 if dim == 2 else (order + 1) * (order + 2) * (order + 3) // 6
        e2d = np.empty((elems.shape[0], nloc * ncomp), dtype=np.int32)
        _check(lib().feel_hip_space_elem2dof(h, e2d))
        self.ctx, self.dim, self.order, self.ncomp = ctx, dim, order, ncomp
        self.coords, self.elems, self.elem2dof = coords, elems, e2d
        self.ndof = int(lib().feel_hip_space_ndof(h))
        self.ne = elems.shape[0]
        return self


class Graph(_Handle):
    _destroy = "feel_hip_graph_destroy"

    def __init__(self, space, ia=None, ja=None):
        if ia is None:
            ia, ja = graph(space.elem2dof, space.ndof)
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        h = _vp()
        _check(lib().feel_hip_graph_create(space.h, space.ndof, ia, ja, C.byref(h)))
        super().__init__(h, keep=(space,))
        self.space, self.ia, self.ja = space, ia, ja
        self.nnz = int(ja.shape[0])

    @classmethod
    def from_space(cls, space):
        """The space's COUPLED pattern built on the device; ia/ja copied back."""
        h = _vp()
        _check(lib().feel_hip_graph_create_from_space(space.h, C.byref(h)))
        self = cls.__new__(cls)
        _Handle.__init__(self, h, keep=(space,))
        nnz = int(lib().feel_hip_graph_nnz(h))
        ia = np.empty(space.ndof + 1, dtype=np.int32)
        ja = np.empty(nnz, dtype=np.int32)
        _check(lib().feel_hip_graph_pattern(h, ia, ja))
        self.space, self.ia, self.ja, self.nnz = space, ia, ja, nnz
        return self

    def spmv_shape(self, precision):
        """(shape index, block, grid) of the SpMV launch for FP64 or FP32 solves; -1 until tuned."""
        sh, bl, gr = C.c_int(), C.c_int(), C.c_int()
        _check(lib().feel_hip_graph_spmv_shape(self.h, precision, C.byref(sh), C.byref(bl), C.byref(gr)))
        return sh.value, bl.value, gr.value

    def set_spmv_shape(self, precision, shape):
        """Pins the SpMV launch shape (bitwise reproducible solves across processes); -1 re-tunes."""
        _check(lib().feel_hip_graph_set_spmv_shape(self.h, precision, shape))


class Matrix(_Handle):
    _destroy = "feel_hip_mat_destroy"

    def __init__(self, graph, precision=FP64):
        h = _vp()
        _check(lib().feel_hip_mat_create(graph.h, precision, C.byref(h)))
        super().__init__(h, keep=(graph,))
        self.graph, self.precision = graph, precision

    def zero(self):
        _check(lib().feel_hip_mat_zero(self.h))

    def values(self, precision=FP64):
        if precision == FP64:
            out = np.zeros(self.graph.nnz)
        else:
            out = np.zeros(self.graph.nnz, np.float32)
        _check(lib().feel_hip_mat_get_values(self.h, precision, out.ctypes.data_as(_vp)))
        return out

    def set_values(self, vals):
        _check(lib().feel_hip_mat_set_values(self.h, np.ascontiguousarray(vals, dtype=np.float64)))

    def add_matrix(self, rows, cols, data):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        data = np.ascontiguousarray(data, dtype=np.float64).reshape(len(rows), len(cols))
        _check(lib().feel_hip_mat_add_matrix(self.h, len(rows), rows, len(cols), cols, data))

    def mult(self, x, y):
        _check(lib().feel_hip_mat_mult(self.h, x.h, y.h))

    def scale(self, a):
        _check(lib().feel_hip_mat_scale(self.h, a))

    def axpy(self, a, X, structure=SAME_NONZERO_PATTERN):
        """self += a X (MatrixSparse::addMatrix(a, X, structure))"""
        _check(lib().feel_hip_mat_axpy(self.h, a, X.h, structure))

    def copy_from(self, src):
        _check(lib().feel_hip_mat_copy(self.h, src.h))

    def diagonal(self):
        ctx = self.graph.space.ctx
        d = Vector(ctx, self.graph.space.ndof)
        _check(lib().feel_hip_mat_diagonal(self.h, d.h))
        return d.get()

    def transpose_into(self, At):
        _check(lib().feel_hip_mat_transpose(self.h, At.h))

    def norm(self, kind):
        out = C.c_double()
        _check(lib().feel_hip_mat_norm(self.h, kind, C.byref(out)))
        return out.value

    def energy(self, v, u, transpose=False):
        """v^T A u (v^T A^T u with transpose)"""
        out = C.c_double()
        _check(lib().feel_hip_mat_energy(self.h, v.h, u.h, int(transpose), C.byref(out)))
        return out.value

    def set_matrix(self, rows, cols, data):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        data = np.ascontiguousarray(data, dtype=np.float64).reshape(len(rows), len(cols))
        _check(lib().feel_hip_mat_set_matrix(self.h, len(rows), rows, len(cols), cols, data))

    def get_entries(self, rows, cols):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros(len(rows))
        _check(lib().feel_hip_mat_get_entries(self.h, len(rows), rows, cols, out))
        return out

    def assemble(self, kind, quad_order, diff=1.0, kdiff=None, mass=1.0, kmass=None, young=None, poisson=None,
                 precision=0, add=False, matrices=MATRICES_AUTO):
        d = BilinearDesc(kind, quad_order, precision, diff, kdiff.h if kdiff else None, mass,
                         kmass.h if kmass else None, young.h if young else None, poisson.h if poisson else None,
                         matrices)
        _check(lib().feel_hip_assemble_bilinear(self.h, C.byref(d), int(add)))

    def zero_rows(self, rows, values=None, rhs=None, mode=ELIMINATION_SYMMETRIC, diag=1.0):
        _check(lib().feel_hip_mat_zero_rows(self.h, rows.h, values.h if values else None, rhs.h if rhs else None,
                                            mode, diag))

    def solve(self, b, x, rtol=1e-8, atol=1e-50, dtol=1e5, maxit=1000, precision=FP64, inner_rtol=1e-2,
              inner_maxit=100000, max_outer=50, check_every=0, inner_operator=0):
        o = SolveOpts(rtol, atol, dtol, maxit, precision, inner_rtol, inner_maxit, max_outer, check_every,
                      inner_operator)
        out = SolveData()
        _check(lib().feel_hip_solve_pcg(self.h, b.h, x.h, C.byref(o), C.byref(out)))
        return out.as_dict()


class Rows(_Handle):
    _destroy = "feel_hip_rows_destroy"

    def __init__(self, ctx, n, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        h = _vp()
        _check(lib().feel_hip_rows_create(ctx.h, n, rows.shape[0], rows, C.byref(h)))
        super().__init__(h, keep=(ctx,))
        self.rows = rows


FN_ZERO, FN_LINEAR, FN_SINSIN = 0, 1, 2


class Function(C.Structure):
    _fields_ = [("kind", C.c_int), ("c", C.c_double * 4)]


def _functions(space, funcs):
    if funcs is None:
        return None
    arr = (Function * space.ncomp)()
    for i, (kind, params) in enumerate(funcs):
        arr[i].kind = kind
        for k, v in enumerate(list(params)[:4]):
            arr[i].c[k] = v
    return arr


def norm_l2(space, u, funcs=None, quad_order=-1):
    """normL2(elements(mesh), u - f): funcs gives (kind, [c0..c3]) per component (None: f = 0)."""
    out = C.c_double()
    _check(lib().feel_hip_norm_l2(space.h, u.h, _functions(space, funcs), quad_order, C.byref(out)))
    return out.value


def norm_semi_h1(space, u, funcs=None, quad_order=-1):
    """H1 seminorm of u - f over elements(mesh)."""
    out = C.c_double()
    _check(lib().feel_hip_norm_semi_h1(space.h, u.h, _functions(space, funcs), quad_order, C.byref(out)))
    return out.value


def assemble_linear(space, b, kind, quad_order, params=(1.0,), add=False):
    p = (C.c_double * 4)(*(list(params) + [0.0] * (4 - len(params))))
    d = LinearDesc(kind, quad_order, p)
    _check(lib().feel_hip_assemble_linear(space.h, C.byref(d), b.h, int(add)))


class Batch(_Handle):
    """Batched / continuously batched solves of affine families A(mu) = sum_t mu_t A_t."""
    _destroy = "feel_hip_batch_destroy"

    def __init__(self, ctx, rtol=1e-8, inner_rtol=1e-2, precision=FP32, hbm_budget=0, use_graph=True,
                 maxit=100000, max_outer=50, group_spmv=None, backfill=0):
        o = BatchOpts()
        lib().feel_hip_batch_opts_default(C.byref(o))
        o.rtol, o.inner_rtol, o.precision, o.hbm_budget, o.use_graph = rtol, inner_rtol, precision, hbm_budget, int(
            use_graph)
        if group_spmv is not None:
            o.group_spmv = int(group_spmv)
        o.maxit, o.max_outer, o.backfill = maxit, max_outer, backfill
        h = _vp()
        _check(lib().feel_hip_batch_create(ctx.h, C.byref(o), C.byref(h)))
        super().__init__(h, keep=(ctx,))
        self.families = []

    def add_family(self, terms, rhs, rows=None):
        arr = (_vp * len(terms))(*[t.h for t in terms])
        fam = C.c_int()
        _check(lib().feel_hip_batch_add_family(self.h, len(terms), arr, rhs.h, rows.h if rows else None, C.byref(fam)))
        self.families.append((terms, rhs, rows))
        return fam.value

    def submit(self, family, mu):
        rid = C.c_int64()
        _check(lib().feel_hip_batch_submit(self.h, family, np.ascontiguousarray(mu, dtype=np.float64), C.byref(rid)))
        return rid.value

    def step(self, iterations=32):
        a, d = C.c_int(), C.c_int()
        _check(lib().feel_hip_batch_step(self.h, iterations, C.byref(a), C.byref(d)))
        return a.value, d.value

    def poll(self):
        rid, sd = C.c_int64(), SolveData()
        got = _check(lib().feel_hip_batch_poll(self.h, C.byref(rid), C.byref(sd)))
        return (rid.value, sd.as_dict()) if got else None

    def solution(self, request, n):
        x = np.zeros(n)
        _check(lib().feel_hip_batch_solution(self.h, request, x.ctypes.data_as(_vp)))
        return x

    def device_bytes(self):
        return lib().feel_hip_batch_device_bytes(self.h)

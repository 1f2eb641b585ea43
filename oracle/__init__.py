"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the CPU baseline and
never as the thing measured or shipped.  The product (feelpp_amd) never
imports this package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_mesh_square.argtypes = [C.c_int, C.c_double, C.c_uint64, _f64p, _i32p]
        L.orc_mesh_kuhn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_uint64, _f64p, _i32p]
        L.orc_lognormal.argtypes = [C.c_long, C.c_double, C.c_double, C.c_uint64, _f64p]
        L.orc_uniform.argtypes = [C.c_long, C.c_double, C.c_double, C.c_int, C.c_uint64, _f64p]
        L.orc_quadrature.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_lagrange_nodes.argtypes = [C.c_int, C.c_int, C.c_void_p]
        L.orc_set_p4_face_weight.restype = None
        L.orc_set_p4_face_weight.argtypes = [C.c_double]
        L.orc_lagrange_eval.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.orc_doftable.restype = C.c_long
        L.orc_doftable.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p]
        L.orc_graph.restype = C.c_long
        L.orc_graph.argtypes = [C.c_int, C.c_int, C.c_long, _i32p, _i32p, C.c_void_p]
        L.orc_assemble_scalar.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p,
                                          C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                          _i32p, _i32p, _f64p]
        L.orc_assemble_scalar_mt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p,
                                             C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_long,
                                             _i32p, _i32p, _f64p, C.c_int]
        L.orc_pcg_mt.argtypes = [C.c_long, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]
        L.orc_norm.restype = C.c_double
        L.orc_norm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p, _f64p, _i32p,
                               _f64p, C.c_int]
        L.orc_assemble_elasticity.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p,
                                              _f64p, _f64p, _i32p, _i32p, _f64p]
        L.orc_assemble_rhs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p,
                                       _i32p, C.c_int, _f64p, _f64p]
        L.orc_boundary_dofs.restype = C.c_long
        L.orc_boundary_dofs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p, _i32p,
                                        C.c_long, C.c_int, _u8p]
        L.orc_dirichlet.argtypes = [C.c_int, C.c_long, _i32p, _i32p, _f64p, _f64p, _u8p, C.c_void_p,
                                    C.c_double]
        L.orc_spmv.restype = None
        L.orc_spmv.argtypes = [C.c_long, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_pcg.argtypes = [C.c_long, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                              C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                              C.POINTER(C.c_int)]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- inputs
def mesh_square(n, jitter=0.0, seed=0):
    coords = np.zeros(((n + 1) ** 2, 2))
    elems = np.zeros((2 * n * n, 3), np.int32)
    lib().orc_mesh_square(n, jitter, seed, coords, elems)
    return coords, elems


def mesh_kuhn(nx, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    elems = np.zeros((6 * nx * ny * nz, 4), np.int32)
    lib().orc_mesh_kuhn(nx, ny, nz, lx, ly, lz, jitter, seed, coords, elems)
    return coords, elems


def lognormal(n, mu, sigma, seed):
    out = np.zeros(n)
    lib().orc_lognormal(n, mu, sigma, seed, out)
    return out


def uniform(n, lo, hi, seed, log=False):
    out = np.zeros(n)
    lib().orc_uniform(n, lo, hi, int(log), seed, out)
    return out


# ---------------------------------------------------------------- reference element
def quadrature(dim, order):
    n = lib().orc_quadrature(dim, order, None, None)
    if n < 0:
        raise ValueError("no quadrature")
    pts = np.zeros((n, dim))
    w = np.zeros(n)
    lib().orc_quadrature(dim, order, _ptr(pts), _ptr(w))
    return pts, w


def lagrange_nodes(dim, order):
    n = lib().orc_lagrange_nodes(dim, order, None)
    if n < 0:
        raise ValueError("unsupported order")
    nodes = np.zeros((n, dim))
    lib().orc_lagrange_nodes(dim, order, _ptr(nodes))
    return nodes


def set_p4_face_weight(w):
    """Replace the P4 face-interior node weight (0: back to warp & blend); a test hook."""
    lib().orc_set_p4_face_weight(float(w))


def lagrange_eval(dim, order, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    npts = pts.shape[0]
    n = lib().orc_lagrange_nodes(dim, order, None)
    phi = np.zeros((n, npts))
    dphi = np.zeros((n, npts, dim))
    lib().orc_lagrange_eval(dim, order, npts, pts, phi, dphi)
    return phi, dphi


# ---------------------------------------------------------------- space / graph
def doftable(dim, order, ncomp, elems):
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    ne = elems.shape[0]
    nl = (order + 1) * (order + 2) // 2 if dim == 2 else (order + 1) * (order + 2) * (order + 3) // 6
    e2d = np.zeros((ne, nl * ncomp), np.int32)
    ndof = lib().orc_doftable(dim, order, ncomp, ne, elems, e2d)
    if ndof < 0:
        raise ValueError("unsupported space")
    return e2d, int(ndof)


def graph(elem2dof, ndof):
    e2d = np.ascontiguousarray(elem2dof, dtype=np.int32)
    ia = np.zeros(ndof + 1, np.int32)
    nnz = lib().orc_graph(e2d.shape[1], e2d.shape[0], ndof, e2d, ia, None)
    ja = np.zeros(nnz, np.int32)
    lib().orc_graph(e2d.shape[1], e2d.shape[0], ndof, e2d, ia, _ptr(ja))
    return ia, ja


# ---------------------------------------------------------------- assembly
def assemble_scalar(dim, order, qorder, coords, elems, elem2dof, ia, ja, diff=1.0, kdiff=None,
                    mass=0.0, kmass=None):
    val = np.zeros(ja.shape[0])
    kd = None if kdiff is None else np.ascontiguousarray(kdiff, dtype=np.float64)
    km = None if kmass is None else np.ascontiguousarray(kmass, dtype=np.float64)
    rc = lib().orc_assemble_scalar(dim, order, qorder, elems.shape[0], np.ascontiguousarray(coords),
                                   np.ascontiguousarray(elems, dtype=np.int32),
                                   np.ascontiguousarray(elem2dof, dtype=np.int32), diff, _ptr(kd), mass,
                                   _ptr(km), ia, ja, val)
    if rc != 0:
        raise RuntimeError(f"orc_assemble_scalar failed ({rc})")
    return val


def assemble_scalar_mt(dim, order, qorder, coords, elems, elem2dof, ia, ja, diff=1.0, kdiff=None,
                       mass=0.0, kmass=None, threads=1):
    """Threaded CPU baseline of assemble_scalar (row ownership, ghost elements recomputed)."""
    val = np.zeros(ja.shape[0])
    kd = None if kdiff is None else np.ascontiguousarray(kdiff, dtype=np.float64)
    km = None if kmass is None else np.ascontiguousarray(kmass, dtype=np.float64)
    rc = lib().orc_assemble_scalar_mt(dim, order, qorder, elems.shape[0], np.ascontiguousarray(coords),
                                      np.ascontiguousarray(elems, dtype=np.int32),
                                      np.ascontiguousarray(elem2dof, dtype=np.int32), diff, _ptr(kd), mass,
                                      _ptr(km), ia.shape[0] - 1, ia, ja, val, threads)
    if rc != 0:
        raise RuntimeError(f"orc_assemble_scalar_mt failed ({rc})")
    return val


def assemble_elasticity(order, qorder, coords, elems, elem2dof, E, nu, ia, ja):
    val = np.zeros(ja.shape[0])
    rc = lib().orc_assemble_elasticity(order, qorder, elems.shape[0], np.ascontiguousarray(coords),
                                       np.ascontiguousarray(elems, dtype=np.int32),
                                       np.ascontiguousarray(elem2dof, dtype=np.int32),
                                       np.ascontiguousarray(E, dtype=np.float64),
                                       np.ascontiguousarray(nu, dtype=np.float64), ia, ja, val)
    if rc != 0:
        raise RuntimeError(f"orc_assemble_elasticity failed ({rc})")
    return val


FK_CONST, FK_SINSIN, FK_VECTOR = 0, 1, 2


def assemble_rhs(dim, order, ncomp, qorder, coords, elems, elem2dof, ndof, fkind, fparams):
    b = np.zeros(ndof)
    fp = np.ascontiguousarray(np.atleast_1d(np.asarray(fparams, dtype=np.float64)))
    rc = lib().orc_assemble_rhs(dim, order, ncomp, qorder, elems.shape[0], np.ascontiguousarray(coords),
                                np.ascontiguousarray(elems, dtype=np.int32),
                                np.ascontiguousarray(elem2dof, dtype=np.int32), fkind, fp, b)
    if rc != 0:
        raise RuntimeError("orc_assemble_rhs failed")
    return b


def boundary_dofs(dim, order, ncomp, coords, elems, elem2dof, ndof, sel=0):
    isb = np.zeros(ndof, np.uint8)
    lib().orc_boundary_dofs(dim, order, ncomp, elems.shape[0], np.ascontiguousarray(coords),
                            np.ascontiguousarray(elems, dtype=np.int32),
                            np.ascontiguousarray(elem2dof, dtype=np.int32), ndof, sel, isb)
    return isb


def dirichlet(mode, ia, ja, val, b, isbnd, g=None, diag=1.0):
    gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    lib().orc_dirichlet(mode, ia.shape[0] - 1, ia, ja, val, b, isbnd, _ptr(gg), diag)


def spmv(ia, ja, val, x):
    y = np.zeros(ia.shape[0] - 1)
    lib().orc_spmv(ia.shape[0] - 1, ia, ja, val, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


FN_ZERO, FN_LINEAR, FN_SINSIN = 0, 1, 2


def norm(dim, order, ncomp, qorder, coords, elems, elem2dof, u, funcs=None, semi=False):
    """L2 norm (or H1 seminorm) of u - f over all elements; funcs: per component (kind, [c0..c3])."""
    funcs = funcs or [(FN_ZERO, [0.0] * 4)] * ncomp
    kinds = np.array([k for k, _ in funcs], dtype=np.int32)
    params = np.array([list(p) + [0.0] * (4 - len(p)) for _, p in funcs], dtype=np.float64).ravel()
    v = lib().orc_norm(dim, order, ncomp, qorder, elems.shape[0], np.ascontiguousarray(coords),
                       np.ascontiguousarray(elems, dtype=np.int32), np.ascontiguousarray(elem2dof, dtype=np.int32),
                       np.ascontiguousarray(u, dtype=np.float64), kinds, params, int(semi))
    if v < 0:
        raise RuntimeError("orc_norm failed")
    return v


def pcg_mt(ia, ja, val, b, x0=None, rtol=1e-8, atol=1e-50, dtol=1e5, maxit=1000, threads=1):
    """Threaded CPU baseline of pcg (same algorithm, reassociated sums)."""
    n = ia.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    its, res, reason = C.c_int(), C.c_double(), C.c_int()
    lib().orc_pcg_mt(n, ia, ja, val, np.ascontiguousarray(b, dtype=np.float64), x, rtol, atol, dtol, maxit,
                     threads, C.byref(its), C.byref(res), C.byref(reason))
    return x, its.value, res.value, reason.value


def pcg(ia, ja, val, b, x0=None, rtol=1e-8, atol=1e-50, dtol=1e5, maxit=1000):
    n = ia.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    its, res, reason = C.c_int(), C.c_double(), C.c_int()
    lib().orc_pcg(n, ia, ja, val, np.ascontiguousarray(b, dtype=np.float64), x, rtol, atol, dtol, maxit,
                  C.byref(its), C.byref(res), C.byref(reason))
    return x, its.value, res.value, reason.value

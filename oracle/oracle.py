"""ctypes binding of the CPU oracle (oracle/liballs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the library.
The oracle restates the reference's serial CPU assembly; see alls_oracle.h
for what each function follows and how it is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liballs_oracle.so")


class OrMesh(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("npe", C.c_int32), ("n_nodes", C.c_int64), ("n_elems", C.c_int64),
        ("coords", C.c_void_p), ("conn", C.c_void_p), ("types", C.c_void_p), ("uniform_type", C.c_int32),
        ("E", C.c_double), ("E_elem", C.c_void_p), ("nu", C.c_double), ("thickness", C.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, P = C.c_void_p, C.POINTER(C.c_double), C.POINTER
        L.or_ke_q4.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, vp, vp]
        L.or_ke_t3.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, vp, vp]
        L.or_ke_hex20.argtypes = [vp, vp, vp, C.c_double, C.c_double, vp, vp]
        L.or_ke_hex20_f32.argtypes = [vp, vp, vp, C.c_float, C.c_float, vp]
        for f in (L.or_ke_q4, L.or_ke_t3, L.or_ke_hex20, L.or_ke_hex20_f32):
            f.restype = None
        L.or_pattern.argtypes = [P(OrMesh), vp, vp]
        L.or_pattern.restype = C.c_int64
        L.or_pattern_bruteforce.argtypes = [P(OrMesh), vp, vp]
        L.or_pattern_bruteforce.restype = C.c_int64
        L.or_assemble.argtypes = [P(OrMesh), vp, vp, vp, vp]
        L.or_assemble.restype = C.c_int64
        L.or_assemble_dense.argtypes = [P(OrMesh), vp]
        L.or_assemble_dense.restype = None
        L.or_precision.argtypes = [C.c_int64, vp, vp, vp]
        L.or_precision.restype = None
        L.or_color_greedy.argtypes = [P(OrMesh), vp]
        L.or_color_greedy.restype = C.c_int32
        L.or_coloring_conflicts.argtypes = [P(OrMesh), vp]
        L.or_coloring_conflicts.restype = C.c_int64
        # csr_checks.c: full-size CSR invariants (OpenMP)
        L.ck_rows_sorted_unique.argtypes = [C.c_int64, vp, vp, C.c_int64]
        L.ck_rows_sorted_unique.restype = C.c_int64
        L.ck_symmetry.argtypes = [C.c_int64, vp, vp, vp, vp, vp]
        L.ck_symmetry.restype = C.c_double
        L.ck_matvec.argtypes = [C.c_int64, vp, vp, vp, vp, vp]
        L.ck_matvec.restype = None
        L.ck_frobenius.argtypes = [C.c_int64, vp, vp, vp]
        L.ck_frobenius.restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def ke_q4(x, y, E=1.0, nu=0.3, t=1.0):
    x = np.ascontiguousarray(x, float); y = np.ascontiguousarray(y, float)
    ke = np.empty(64); ka = np.empty(64)
    lib().or_ke_q4(_p(x), _p(y), E, nu, t, _p(ke), _p(ka))
    return ke.reshape(8, 8), ka.reshape(8, 8)


def ke_t3(x, y, E=1.0, nu=0.3, t=1.0):
    x = np.ascontiguousarray(x, float); y = np.ascontiguousarray(y, float)
    ke = np.empty(36); ka = np.empty(36)
    lib().or_ke_t3(_p(x), _p(y), E, nu, t, _p(ke), _p(ka))
    return ke.reshape(6, 6), ka.reshape(6, 6)


def ke_hex20(x, y, z, E=1.0, nu=0.3):
    x, y, z = (np.ascontiguousarray(a, float) for a in (x, y, z))
    ke = np.empty(3600); ka = np.empty(3600)
    lib().or_ke_hex20(_p(x), _p(y), _p(z), E, nu, _p(ke), _p(ka))
    return ke.reshape(60, 60), ka.reshape(60, 60)


def ke_hex20_f32(x, y, z, E=1.0, nu=0.3):
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    ke = np.empty(3600, np.float32)
    lib().or_ke_hex20_f32(_p(x), _p(y), _p(z), E, nu, _p(ke))
    return ke.reshape(60, 60)


class OracleMesh:
    """Holds the arrays of a mesh (an alls.MeshArrays or compatible)."""

    def __init__(self, m):
        self.src = m
        self.coords = np.ascontiguousarray(m.coords, np.float64)
        self.conn = np.ascontiguousarray(m.conn, np.int32)
        self.types = None if m.elem_type is None else np.ascontiguousarray(m.elem_type, np.uint8)
        self.E_elem = None if m.E_per_elem is None else np.ascontiguousarray(m.E_per_elem, np.float64)
        s = OrMesh()
        s.dim, s.npe = m.dim, m.npe
        s.n_nodes, s.n_elems = self.coords.size // m.dim, self.conn.size // m.npe
        s.coords, s.conn, s.types = _p(self.coords), _p(self.conn), _p(self.types)
        s.uniform_type, s.E, s.E_elem = m.uniform_type, m.E, _p(self.E_elem)
        s.nu, s.thickness = m.nu, m.thickness
        self.s = s
        self.n_dofs = s.n_nodes * m.dim


def pattern(m):
    om = OracleMesh(m)
    rp = np.empty(om.n_dofs + 1, np.int64)
    nnz = lib().or_pattern(C.byref(om.s), _p(rp), None)
    if nnz < 0:
        raise ValueError("invalid mesh")
    ci = np.empty(max(nnz, 1), np.int32)
    lib().or_pattern(C.byref(om.s), _p(rp), _p(ci))
    return rp, ci[:nnz]


def pattern_bruteforce(m):
    om = OracleMesh(m)
    rp = np.empty(om.n_dofs + 1, np.int64)
    nnz = lib().or_pattern_bruteforce(C.byref(om.s), _p(rp), None)
    if nnz < 0:
        raise ValueError("mesh too large or invalid for brute force")
    ci = np.empty(max(nnz, 1), np.int32)
    lib().or_pattern_bruteforce(C.byref(om.s), _p(rp), _p(ci))
    return rp, ci[:nnz]


def assemble(m, rp=None, ci=None, with_abs=True):
    """Serial element-order assembly (Acc:329-380) -> (rp, ci, val, absval)."""
    if rp is None:
        rp, ci = pattern(m)
    om = OracleMesh(m)
    nnz = int(rp[-1])
    val = np.empty(max(nnz, 1))
    av = np.empty(max(nnz, 1)) if with_abs else None
    miss = lib().or_assemble(C.byref(om.s), _p(rp), _p(ci), _p(val), _p(av))
    if miss:
        raise AssertionError(f"{miss} contributions missing from the pattern")
    return rp, ci, val[:nnz], (av[:nnz] if with_abs else None)


def assemble_dense(m):
    om = OracleMesh(m)
    K = np.empty((om.n_dofs, om.n_dofs))
    lib().or_assemble_dense(C.byref(om.s), _p(K))
    return K


def precision(ref, test):
    """Acc:896-960 metrics: max_abs, max_rel, mean_abs, mean_rel, frobenius_rel."""
    ref = np.ascontiguousarray(ref, float); test = np.ascontiguousarray(test, float)
    out = np.empty(5)
    lib().or_precision(ref.size, _p(ref), _p(test), _p(out))
    return dict(zip(["max_abs", "max_rel", "mean_abs", "mean_rel", "frobenius_rel"], out.tolist()))


def color_greedy(m):
    om = OracleMesh(m)
    col = np.empty(om.s.n_elems, np.int32)
    n = lib().or_color_greedy(C.byref(om.s), _p(col))
    return n, col


def coloring_conflicts(m, colors):
    om = OracleMesh(m)
    colors = np.ascontiguousarray(colors, np.int32)
    return lib().or_coloring_conflicts(C.byref(om.s), _p(colors))


# ----------------------------------------------- full-size CSR invariants ---
# (csr_checks.c; rp must be int64, ci int32, v float64, all contiguous)

def _csr_args(rp, ci, v=None):
    rp = np.ascontiguousarray(rp, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    v = None if v is None else np.ascontiguousarray(v, np.float64)
    return rp, ci, v


def rows_sorted_unique_violations(rp, ci, n_cols=None):
    rp, ci, _ = _csr_args(rp, ci)
    n = rp.size - 1
    return int(lib().ck_rows_sorted_unique(n, _p(rp), _p(ci), n if n_cols is None else n_cols))


def symmetry(rp, ci, v):
    """(max |K_ij - K_ji|, entries without a transpose, max |K_ij|)."""
    rp, ci, v = _csr_args(rp, ci, v)
    miss = C.c_int64(0)
    vmax = C.c_double(0.0)
    d = lib().ck_symmetry(rp.size - 1, _p(rp), _p(ci), _p(v), C.addressof(miss), C.addressof(vmax))
    return float(d), int(miss.value), float(vmax.value)


def matvec(rp, ci, v, x):
    rp, ci, v = _csr_args(rp, ci, v)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(rp.size - 1)
    lib().ck_matvec(rp.size - 1, _p(rp), _p(ci), _p(v), _p(x), _p(y))
    return y


def frobenius_rel(a, b):
    """||a - b|| / ||b||."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.empty(2)
    lib().ck_frobenius(a.size, _p(a), _p(b), _p(out))
    return float(np.sqrt(out[0]) / max(np.sqrt(out[1]), 1e-300))

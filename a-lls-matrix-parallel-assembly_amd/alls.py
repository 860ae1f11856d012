"""ctypes binding of liballs.so (the C ABI in include/alls.h).

This is the Python face of the library used by tests, bench.py and the
smoke check.  It loads the in-tree liballs.so and raises if it is missing:
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liballs.so")

OK, E_INVALID, E_OOM, E_POOL_OVERFLOW, E_HIP, E_BUSY, E_NO_DEVICE, E_UNSUPPORTED = range(8)
CACHE_COST, CACHE_EAGER = 0, 1
Q4, T3, HEX20 = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1
PREC_F64, PREC_F32_MFMA = 0, 1
GEN_Q4, GEN_T3, GEN_MIXED, GEN_HEX20 = 0, 1, 2, 3


class Mesh(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nodes_per_elem", C.c_int32),
        ("n_nodes", C.c_int64), ("n_elems", C.c_int64),
        ("coords", C.c_void_p), ("conn", C.c_void_p), ("elem_type", C.c_void_p),
        ("uniform_type", C.c_int32), ("mem", C.c_int32),
        ("E", C.c_double), ("E_per_elem", C.c_void_p),
        ("nu", C.c_double), ("thickness", C.c_double),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("chunk_slots", C.c_int32), ("block_threads", C.c_int32),
        ("row_ptr_i64", C.c_int32), ("timing", C.c_int32), ("grid_per_cu", C.c_int32),
        ("autotune", C.c_int32), ("val_f32", C.c_int32),
    ]


class Csr(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64), ("nnz", C.c_int64), ("row_ptr_is_i64", C.c_int32), ("pattern_id", C.c_int32),
        ("row_ptr", C.c_void_p), ("col_ind", C.c_void_p), ("val", C.c_void_p),
        ("cap_rows", C.c_int64), ("cap_nnz", C.c_int64), ("val_is_f32", C.c_int32), ("_pad", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_init", C.c_double), ("ms_insert", C.c_double),
        ("ms_count", C.c_double), ("ms_scan", C.c_double), ("ms_write", C.c_double),
        ("n_elems", C.c_int64), ("n_rows", C.c_int64), ("nnz", C.c_int64),
        ("list_entries", C.c_int64), ("pool_chunks_used", C.c_int64), ("pool_chunks_cap", C.c_int64),
        ("scratch_bytes", C.c_int64), ("chunk_slots", C.c_int32), ("retries", C.c_int32),
        ("n_colors", C.c_int32), ("big_rows", C.c_int32), ("graph_launches", C.c_int32), ("_pad", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerCounts(C.Structure):
    _fields_ = [
        ("completed", C.c_int64), ("batched", C.c_int64), ("passed", C.c_int64),
        ("builds", C.c_int64 * 2), ("hits", C.c_int64 * 2), ("evictions", C.c_int64 * 2),
        ("evicted_unused", C.c_int64 * 2),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("n_streams", C.c_int32), ("hbm_budget_bytes", C.c_int64),
        ("pattern_cache_entries", C.c_int32), ("cache_policy", C.c_int32),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
        ("nnz", C.c_int64), ("cache_hit", C.c_int32), ("status", C.c_int32),
        ("batch_size", C.c_int32), ("reserved", C.c_int32),
    ]


class GenSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("jitter", C.c_double), ("seed", C.c_uint64), ("random_diagonal", C.c_int32),
        ("permute_nodes", C.c_int32), ("split_fraction", C.c_double), ("E", C.c_double),
        ("E_sigma", C.c_double), ("nu", C.c_double), ("thickness", C.c_double),
    ]


class HostMesh(C.Structure):
    _fields_ = [("mesh", Mesh), ("structured_color", C.c_void_p), ("n_structured_colors", C.c_int32),
                ("_pad", C.c_int32)]


P = C.POINTER
vp = C.c_void_p

# name -> (restype, argtypes) for every function declared in include/alls.h.
class Peaks(C.Structure):
    _fields_ = [("copy_GBs", C.c_double), ("mfma_f64_TFLOPs", C.c_double), ("mfma_f32_TFLOPs", C.c_double),
                ("atomic_f64_Gops", C.c_double)]


SIGNATURES = {
    "alls_status_str": (C.c_char_p, [C.c_int]),
    "alls_ctx_create": (C.c_int, [P(vp), C.c_int, vp]),
    "alls_ctx_destroy": (None, [vp]),
    "alls_ctx_last_error": (C.c_char_p, [vp]),
    "alls_last_stats": (C.c_int, [vp, P(Stats)]),
    "alls_device_count": (C.c_int, []),
    "alls_assemble": (C.c_int, [vp, P(Mesh), P(Opts), P(Csr)]),
    "alls_assemble_noalloc": (C.c_int, [vp, P(Mesh), P(Opts), P(Csr)]),
    "alls_csr_download": (C.c_int, [vp, P(Csr), vp, vp, vp]),
    "alls_csr_free": (None, [vp, P(Csr)]),
    "alls_pattern_create": (C.c_int, [vp, P(Mesh), P(Opts), P(vp), P(Csr)]),
    "alls_assemble_values": (C.c_int, [vp, vp, vp, C.c_int32, P(Csr)]),
    "alls_pattern_nnz": (C.c_int64, [vp]),
    "alls_pattern_free": (None, [vp, vp]),
    "alls_color": (C.c_int, [vp, vp, P(vp)]),
    "alls_coloring_from_host": (C.c_int, [vp, vp, vp, P(vp)]),
    "alls_coloring_count": (C.c_int32, [vp]),
    "alls_coloring_download": (C.c_int, [vp, vp, vp]),
    "alls_coloring_free": (None, [vp, vp]),
    "alls_assemble_colored": (C.c_int, [vp, vp, vp, vp, C.c_int32, P(Csr)]),
    "alls_assemble_batch": (C.c_int, [vp, C.c_int32, P(Mesh), P(Opts), P(Csr), P(C.c_int64), P(C.c_int64)]),
    "alls_autotune": (C.c_int, [vp, P(Mesh), P(Opts), P(Opts)]),
    "alls_device_peaks": (C.c_int, [vp, P(Peaks)]),
    "alls_csr_write_mtx": (C.c_int, [vp, P(Csr), C.c_char_p]),
    "alls_mesh_save": (C.c_int, [P(Mesh), C.c_char_p]),
    "alls_mesh_load": (C.c_int, [C.c_char_p, P(HostMesh)]),
    "alls_server_create": (C.c_int, [P(ServerConfig), P(vp)]),
    "alls_server_submit": (C.c_int, [vp, P(Mesh), C.c_uint64, P(Opts), P(C.c_int64)]),
    "alls_server_wait": (C.c_int, [vp, C.c_int64, P(Csr), P(RequestStats)]),
    "alls_server_release": (None, [vp, P(Csr)]),
    "alls_server_counters": (C.c_int, [vp, P(ServerCounts)]),
    "alls_server_destroy": (None, [vp]),
    "alls_gen_mesh": (C.c_int, [P(GenSpec), P(HostMesh)]),
    "alls_host_mesh_free": (None, [P(HostMesh)]),
}

_lib = None


def lib():
    """Loads liballs.so (built in-tree); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class AllsError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{lib().alls_status_str(status).decode()} ({status}): {msg}")


def check(status, ctx=None):
    if status != OK:
        msg = ""
        if ctx is not None:
            m = lib().alls_ctx_last_error(ctx.handle)
            msg = m.decode() if m else ""
        raise AllsError(status, msg)


# ----------------------------------------------------------- mesh handling --
class MeshArrays:
    """Owns numpy arrays of a mesh and exposes an alls_mesh view of them."""

    def __init__(self, coords, conn, dim, nodes_per_elem, uniform_type=Q4, elem_type=None,
                 E=1.0, E_per_elem=None, nu=0.3, thickness=1.0, structured_color=None,
                 n_structured_colors=0):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1)
        self.dim = int(dim)
        self.npe = int(nodes_per_elem)
        self.uniform_type = int(uniform_type)
        self.elem_type = None if elem_type is None else np.ascontiguousarray(elem_type, dtype=np.uint8)
        self.E = float(E)
        self.E_per_elem = None if E_per_elem is None else np.ascontiguousarray(E_per_elem, dtype=np.float64)
        self.nu = float(nu)
        self.thickness = float(thickness)
        self.structured_color = structured_color
        self.n_structured_colors = n_structured_colors

    @property
    def n_nodes(self):
        return self.coords.size // self.dim

    @property
    def n_elems(self):
        return self.conn.size // self.npe

    @property
    def n_dofs(self):
        return self.n_nodes * self.dim

    def c_mesh(self):
        m = Mesh()
        m.dim, m.nodes_per_elem = self.dim, self.npe
        m.n_nodes, m.n_elems = self.n_nodes, self.n_elems
        m.coords = self.coords.ctypes.data
        m.conn = self.conn.ctypes.data
        m.elem_type = self.elem_type.ctypes.data if self.elem_type is not None else None
        m.uniform_type, m.mem = self.uniform_type, MEM_HOST
        m.E = self.E
        m.E_per_elem = self.E_per_elem.ctypes.data if self.E_per_elem is not None else None
        m.nu, m.thickness = self.nu, self.thickness
        return m

    def with_E(self, E_per_elem):
        return MeshArrays(self.coords, self.conn, self.dim, self.npe, self.uniform_type, self.elem_type,
                          self.E, E_per_elem, self.nu, self.thickness, self.structured_color,
                          self.n_structured_colors)


def generate(kind, nx, ny=None, nz=0, jitter=0.0, seed=0, random_diagonal=0, permute_nodes=0,
             split_fraction=0.5, E=1.0, E_sigma=0.0, nu=0.3, thickness=1.0):
    """Synthetic benchmark mesh from the library's generator (host only)."""
    s = GenSpec()
    s.kind, s.nx, s.ny, s.nz = kind, nx, ny if ny is not None else nx, nz
    s.jitter, s.seed = jitter, seed
    s.random_diagonal, s.permute_nodes, s.split_fraction = random_diagonal, permute_nodes, split_fraction
    s.E, s.E_sigma, s.nu, s.thickness = E, E_sigma, nu, thickness
    h = HostMesh()
    check(lib().alls_gen_mesh(C.byref(s), C.byref(h)))
    return _from_host_mesh(h)


def load_mesh(path):
    """alls_mesh_load: a mesh saved with save_mesh (binary file)."""
    h = HostMesh()
    check(lib().alls_mesh_load(os.fsencode(path), C.byref(h)))
    return _from_host_mesh(h)


def save_mesh(mesh, path):
    """alls_mesh_save: writes a host mesh to a binary file."""
    m = mesh.c_mesh()
    check(lib().alls_mesh_save(C.byref(m), os.fsencode(path)))


def _from_host_mesh(h):
    """Copies a library-owned alls_host_mesh into numpy arrays and frees it."""
    try:
        m = h.mesh
        dim, npe = m.dim, m.nodes_per_elem
        nn, ne = m.n_nodes, m.n_elems

        def arr(ptr, ctype, n):
            if not ptr or n == 0:
                return None
            return np.ctypeslib.as_array(C.cast(ptr, P(ctype)), shape=(n,)).copy()

        coords = arr(m.coords, C.c_double, nn * dim)
        conn = arr(m.conn, C.c_int32, ne * npe)
        et = arr(m.elem_type, C.c_uint8, ne)
        Ee = arr(m.E_per_elem, C.c_double, ne)
        col = arr(h.structured_color, C.c_int32, ne)
        return MeshArrays(coords, conn, dim, npe, m.uniform_type, et, m.E, Ee, m.nu, m.thickness, col,
                          h.n_structured_colors)
    finally:
        lib().alls_host_mesh_free(C.byref(h))


# ---------------------------------------------------------------- context --
class Context:
    def __init__(self, device=0, stream=None):
        h = vp()
        check(lib().alls_ctx_create(C.byref(h), device, stream))
        self.handle = h

    def close(self):
        if self.handle:
            lib().alls_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        check(lib().alls_last_stats(self.handle, C.byref(s)), self)
        return s.as_dict()

    def assemble(self, mesh, precision=PREC_F64, chunk_slots=0, block_threads=0, row_ptr_i64=0,
                 timing=1, out=None, c_mesh=None, grid_per_cu=0, autotune=0, val_f32=0):
        """Plain A-LLS assembly; returns a DeviceCsr (reused if `out` is given).
        val_f32=1 (Hex20): float output values, downloaded widened to float64."""
        o = Opts()
        o.precision, o.chunk_slots, o.block_threads = precision, chunk_slots, block_threads
        o.row_ptr_i64, o.timing = row_ptr_i64, timing
        o.grid_per_cu, o.autotune, o.val_f32 = grid_per_cu, autotune, val_f32
        out = out or DeviceCsr(self)
        m = c_mesh if c_mesh is not None else mesh.c_mesh()
        check(lib().alls_assemble(self.handle, C.byref(m), C.byref(o), C.byref(out.csr)), self)
        return out


    def assemble_noalloc(self, dmesh, out, precision=PREC_F64, chunk_slots=0, block_threads=0, grid_per_cu=0,
                         c_mesh=None, val_f32=0):
        """alls_assemble_noalloc: a DeviceMesh into the caller's CallerCsr
        buffers, with no device allocation by the library."""
        o = Opts()
        o.precision, o.chunk_slots, o.block_threads = precision, chunk_slots, block_threads
        o.grid_per_cu, o.timing, o.val_f32 = grid_per_cu, 1, val_f32
        m = c_mesh if c_mesh is not None else dmesh.c_mesh()
        check(lib().alls_assemble_noalloc(self.handle, C.byref(m), C.byref(o), C.byref(out.csr)), self)
        return out

    def assemble_batch(self, meshes, precision=PREC_F64, chunk_slots=0, block_threads=0, grid_per_cu=0,
                       autotune=0, out=None, c_meshes=None, val_f32=0):
        """alls_assemble_batch: one pipeline over all meshes (same dim).  Returns
        (DeviceCsr of the block-diagonal result, row_off, nnz_off), offsets as
        numpy int64 arrays of len(meshes)+1."""
        n = len(meshes) if c_meshes is None else len(c_meshes)
        arr = (Mesh * n)(*[(c_meshes[k] if c_meshes is not None else meshes[k].c_mesh()) for k in range(n)])
        o = Opts()
        o.precision, o.chunk_slots, o.block_threads = precision, chunk_slots, block_threads
        o.grid_per_cu, o.autotune, o.timing, o.val_f32 = grid_per_cu, autotune, 1, val_f32
        out = out or DeviceCsr(self)
        row_off = np.zeros(n + 1, np.int64)
        nnz_off = np.zeros(n + 1, np.int64)
        check(lib().alls_assemble_batch(self.handle, n, arr, C.byref(o), C.byref(out.csr),
                                        row_off.ctypes.data_as(P(C.c_int64)), nnz_off.ctypes.data_as(P(C.c_int64))),
              self)
        return out, row_off, nnz_off

    def device_peaks(self):
        """alls_device_peaks: measured copy, MFMA and FP64-atomic rates of this device."""
        pk = Peaks()
        check(lib().alls_device_peaks(self.handle, C.byref(pk)), self)
        return {k: getattr(pk, k) for k, _ in Peaks._fields_}

    def autotune(self, mesh, precision=PREC_F64, c_mesh=None):
        """alls_autotune: the fastest launch setting for this mesh's kind and size class."""
        base, best = Opts(), Opts()
        base.precision, base.timing = precision, 1
        m = c_mesh if c_mesh is not None else mesh.c_mesh()
        check(lib().alls_autotune(self.handle, C.byref(m), C.byref(base), C.byref(best)), self)
        return {"block_threads": best.block_threads, "chunk_slots": best.chunk_slots,
                "grid_per_cu": best.grid_per_cu}


def split_batch(csr, row_off, nnz_off):
    """Per-mesh CSR triples (local row/column ids) from a downloaded batch result."""
    rp, ci, v = csr
    res = []
    for k in range(len(row_off) - 1):
        r0, r1, z0, z1 = row_off[k], row_off[k + 1], nnz_off[k], nnz_off[k + 1]
        res.append((rp[r0:r1 + 1] - z0, ci[z0:z1] - r0, v[z0:z1]))
    return res


class Pattern:
    """Cached sparsity pattern + scatter map (alls_pattern_create)."""

    def __init__(self, ctx, mesh, precision=PREC_F64, chunk_slots=0, out=None, c_mesh=None):
        self.ctx = ctx
        o = Opts()
        o.precision, o.chunk_slots, o.timing = precision, chunk_slots, 1
        self.out = out or DeviceCsr(ctx)
        h = vp()
        m = c_mesh if c_mesh is not None else mesh.c_mesh()
        check(lib().alls_pattern_create(ctx.handle, C.byref(m), C.byref(o), C.byref(h), C.byref(self.out.csr)), ctx)
        self.handle = h
        self.mesh = mesh

    @property
    def nnz(self):
        return lib().alls_pattern_nnz(self.handle)

    def values(self, E_per_elem=None, out=None, device_ptr=None):
        """Value-only re-assembly with a new E field (numpy array or device pointer)."""
        out = out or DeviceCsr(self.ctx)
        if device_ptr is not None:
            check(lib().alls_assemble_values(self.ctx.handle, self.handle, device_ptr, MEM_DEVICE,
                                             C.byref(out.csr)), self.ctx)
        else:
            E = None if E_per_elem is None else np.ascontiguousarray(E_per_elem, np.float64)
            check(lib().alls_assemble_values(self.ctx.handle, self.handle, None if E is None else E.ctypes.data,
                                             MEM_HOST, C.byref(out.csr)), self.ctx)
        return out

    def color_gpu(self):
        h = vp()
        check(lib().alls_color(self.ctx.handle, self.handle, C.byref(h)), self.ctx)
        return Coloring(self.ctx, h)

    def color_from_host(self, colors):
        colors = np.ascontiguousarray(colors, np.int32)
        h = vp()
        check(lib().alls_coloring_from_host(self.ctx.handle, self.handle, colors.ctypes.data, C.byref(h)), self.ctx)
        return Coloring(self.ctx, h)

    def colored(self, coloring, E_per_elem=None, out=None, device_ptr=None):
        out = out or DeviceCsr(self.ctx)
        if device_ptr is not None:
            check(lib().alls_assemble_colored(self.ctx.handle, self.handle, coloring.handle, device_ptr, MEM_DEVICE,
                                              C.byref(out.csr)), self.ctx)
        else:
            E = None if E_per_elem is None else np.ascontiguousarray(E_per_elem, np.float64)
            check(lib().alls_assemble_colored(self.ctx.handle, self.handle, coloring.handle,
                                              None if E is None else E.ctypes.data, MEM_HOST, C.byref(out.csr)),
                  self.ctx)
        return out

    def free(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().alls_pattern_free(self.ctx.handle, self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Coloring:
    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle

    @property
    def n_colors(self):
        return lib().alls_coloring_count(self.handle)

    def download(self, n_elems):
        c = np.empty(n_elems, np.int32)
        check(lib().alls_coloring_download(self.ctx.handle, self.handle, c.ctypes.data), self.ctx)
        return c

    def free(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            lib().alls_coloring_free(self.ctx.handle, self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCsr:
    def __init__(self, ctx):
        self.ctx = ctx
        self.csr = Csr()

    @property
    def nnz(self):
        return self.csr.nnz

    @property
    def n_rows(self):
        return self.csr.n_rows

    def write_mtx(self, path):
        """alls_csr_write_mtx: Matrix Market coordinate file."""
        check(lib().alls_csr_write_mtx(self.ctx.handle, C.byref(self.csr), os.fsencode(path)), self.ctx)

    def download(self):
        """(row_ptr int64, col_ind int32, val float64) as numpy arrays."""
        c = self.csr
        rp = np.empty(c.n_rows + 1, dtype=np.int64)
        ci = np.empty(max(c.nnz, 1), dtype=np.int32)
        v = np.empty(max(c.nnz, 1), dtype=np.float64)
        check(lib().alls_csr_download(self.ctx.handle, C.byref(c), rp.ctypes.data, ci.ctypes.data,
                                      v.ctypes.data), self.ctx)
        return rp, ci[:c.nnz], v[:c.nnz]

    def free(self):
        if self.ctx is not None and self.ctx.handle:
            lib().alls_csr_free(self.ctx.handle, C.byref(self.csr))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CallerCsr:
    """Device CSR buffers owned by the caller (for alls_assemble_noalloc):
    row_ptr of cap_rows entries (int64 if row_ptr_i64), col_ind and val of
    cap_nnz entries, allocated here with hipMalloc and freed with hipFree."""

    def __init__(self, ctx, cap_rows, cap_nnz, row_ptr_i64=0, val_f32=0):
        self.ctx = ctx
        h = hip()
        self.csr = Csr()
        self.csr.row_ptr = h.malloc(cap_rows * (8 if row_ptr_i64 else 4))
        self.csr.col_ind = h.malloc(max(cap_nnz, 1) * 4)
        self.csr.val = h.malloc(max(cap_nnz, 1) * (4 if val_f32 else 8))
        self.csr.cap_rows, self.csr.cap_nnz = cap_rows, cap_nnz
        self.csr.row_ptr_is_i64, self.csr.val_is_f32 = row_ptr_i64, val_f32

    def download(self):
        return DeviceCsr.download(self)

    def free(self):
        h = hip()
        for p in (self.csr.row_ptr, self.csr.col_ind, self.csr.val):
            h.free(p)
        self.csr = Csr()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------- device-resident inputs --
class _Hip:
    """Minimal HIP runtime access (device buffers for device-resident meshes)."""

    def __init__(self):
        lib()  # loads libamdhip64 as a dependency of liballs
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        self.L.hipMalloc.argtypes = [P(vp), C.c_size_t]
        self.L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [vp]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        self.L.hipMemGetInfo.argtypes = [P(C.c_size_t), P(C.c_size_t)]

    def mem_info(self):
        """(free, total) bytes of the current device's HBM."""
        f, t = C.c_size_t(), C.c_size_t()
        r = self.L.hipMemGetInfo(C.byref(f), C.byref(t))
        if r != 0:
            raise RuntimeError(f"hipMemGetInfo failed: {r}")
        return f.value, t.value

    def malloc(self, n):
        p = vp()
        r = self.L.hipMalloc(C.byref(p), max(int(n), 1))
        if r != 0:
            raise MemoryError(f"hipMalloc({n}) failed: {r}")
        return p.value

    def htod(self, dst, arr):
        if arr.nbytes:
            r = self.L.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1)
            if r != 0:
                raise RuntimeError(f"hipMemcpy failed: {r}")

    def dtod(self, dst, src, n):
        r = self.L.hipMemcpy(dst, src, int(n), 3)   # hipMemcpyDeviceToDevice
        if r != 0:
            raise RuntimeError(f"hipMemcpy failed: {r}")

    def free(self, p):
        if p:
            self.L.hipFree(p)

    def sync(self):
        return self.L.hipDeviceSynchronize()


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = _Hip()
    return _hip


class DeviceMesh:
    """Copies a MeshArrays to device memory once; c_mesh() has mem=DEVICE."""

    def __init__(self, mesh):
        self.mesh = mesh
        h = hip()
        self.bufs = []

        def up(a):
            if a is None:
                return None
            p = h.malloc(a.nbytes)
            h.htod(p, a)
            self.bufs.append(p)
            return p

        self.coords = up(mesh.coords)
        self.conn = up(mesh.conn)
        self.elem_type = up(mesh.elem_type)
        self.E_per_elem = up(mesh.E_per_elem)

    def c_mesh(self, E_per_elem_ptr=None):
        m = self.mesh.c_mesh()
        m.coords, m.conn, m.elem_type = self.coords, self.conn, self.elem_type
        m.E_per_elem = E_per_elem_ptr if E_per_elem_ptr is not None else self.E_per_elem
        m.mem = MEM_DEVICE
        return m

    def free(self):
        for p in self.bufs:
            hip().free(p)
        self.bufs = []

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------------- server --
class Server:
    """Continuous-batching assembly service (alls_server_*)."""

    def __init__(self, device=0, n_streams=3, hbm_budget_bytes=0, pattern_cache_entries=0, cache_policy=CACHE_COST):
        cfg = ServerConfig()
        cfg.device, cfg.n_streams = device, n_streams
        cfg.hbm_budget_bytes, cfg.pattern_cache_entries = hbm_budget_bytes, pattern_cache_entries
        cfg.cache_policy = cache_policy
        h = vp()
        check(lib().alls_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, mesh, key=0, precision=PREC_F64, c_mesh=None):
        """Enqueue a request; `mesh` (its arrays) is kept alive until wait()."""
        o = Opts()
        o.precision = precision
        m = c_mesh if c_mesh is not None else mesh.c_mesh()
        t = C.c_int64()
        check(lib().alls_server_submit(self.handle, C.byref(m), C.c_uint64(key), C.byref(o), C.byref(t)))
        self._keep[t.value] = (mesh, m)
        return t.value

    def wait(self, ticket, keep_output=True):
        """Returns (Csr or None, stats dict); the Csr must be released()."""
        out = Csr()
        st = RequestStats()
        rc = lib().alls_server_wait(self.handle, ticket, C.byref(out) if keep_output else None, C.byref(st))
        self._keep.pop(ticket, None)
        stats = {k: getattr(st, k) for k, _ in RequestStats._fields_}
        if rc != OK:
            raise AllsError(rc, f"request {ticket}")
        return (out if keep_output else None), stats

    def download(self, csr):
        """Host copy of a server-owned CSR (uses a temporary context)."""
        rp = np.empty(csr.n_rows + 1, dtype=np.int64)
        ci = np.empty(max(csr.nnz, 1), dtype=np.int32)
        v = np.empty(max(csr.nnz, 1), dtype=np.float64)
        with Context(0) as ctx:
            check(lib().alls_csr_download(ctx.handle, C.byref(csr), rp.ctypes.data, ci.ctypes.data, v.ctypes.data))
        return rp, ci[:csr.nnz], v[:csr.nnz]

    def release(self, csr):
        lib().alls_server_release(self.handle, C.byref(csr))

    def counters(self):
        """alls_server_counters: what the server has done ([2D, Hex20] pairs)."""
        c = ServerCounts()
        check(lib().alls_server_counters(self.handle, C.byref(c)))
        out = {}
        for k, t in ServerCounts._fields_:
            v = getattr(c, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out

    def close(self):
        if self.handle:
            lib().alls_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

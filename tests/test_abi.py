"""C-ABI boundary checks that need no GPU: the library loads, exports every
function include/alls.h declares, and its host-side pieces (status codes,
mesh generators, context creation without a device) behave."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import alls

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "alls.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(alls_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(alls.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    assert set(alls.SIGNATURES) == set(names)


def test_status_strings():
    for s in range(8):
        assert alls.lib().alls_status_str(s)
    assert b"unknown" in alls.lib().alls_status_str(99)


def test_struct_layouts_match_the_c_header(tmp_path):
    # Compile a probe against include/alls.h and compare sizes and offsets
    # with the ctypes mirrors.
    structs = {"alls_mesh": alls.Mesh, "alls_opts": alls.Opts, "alls_csr": alls.Csr,
               "alls_stats": alls.Stats, "alls_gen_spec": alls.GenSpec, "alls_host_mesh": alls.HostMesh,
               "alls_server_config": alls.ServerConfig, "alls_request_stats": alls.RequestStats,
               "alls_peaks": alls.Peaks, "alls_server_counts": alls.ServerCounts}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "alls.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


@pytest.mark.skipif(alls.lib().alls_device_count() > 0, reason="host without a GPU only")
def test_ctx_create_without_device_reports_no_device():
    h = ctypes.c_void_p()
    assert alls.lib().alls_ctx_create(ctypes.byref(h), 0, None) == alls.E_NO_DEVICE
    assert not h.value


def test_null_arguments_are_rejected():
    assert alls.lib().alls_assemble(None, None, None, None) == alls.E_INVALID
    assert alls.lib().alls_csr_download(None, None, None, None, None) == alls.E_INVALID


def test_generator_sizes_and_determinism():
    m = alls.generate(alls.GEN_Q4, 20, 10, jitter=0.2, seed=3, E_sigma=0.5)
    assert (m.n_nodes, m.n_elems, m.npe) == (21 * 11, 200, 4)
    m2 = alls.generate(alls.GEN_Q4, 20, 10, jitter=0.2, seed=3, E_sigma=0.5)
    assert np.array_equal(m.coords, m2.coords) and np.array_equal(m.E_per_elem, m2.E_per_elem)
    m3 = alls.generate(alls.GEN_Q4, 20, 10, jitter=0.2, seed=4, E_sigma=0.5)
    assert not np.array_equal(m.coords, m3.coords)
    # boundary nodes stay on the unit square; jitter is bounded by 0.2h
    X = m.coords.reshape(-1, 2)
    grid = np.stack(np.meshgrid(np.arange(21) / 20, np.arange(11) / 10), -1).reshape(-1, 2)
    assert np.abs(X - grid).max() <= 0.2 / 10 + 1e-15
    # lognormal(0, 0.5): log E has mean ~0 and std ~0.5
    m4 = alls.generate(alls.GEN_Q4, 200, 200, seed=1, E_sigma=0.5)
    le = np.log(m4.E_per_elem)
    assert abs(le.mean()) < 0.01 and abs(le.std() - 0.5) < 0.01


def test_generator_t3_mixed_hex20():
    t3 = alls.generate(alls.GEN_T3, 10, 8, jitter=0.3, seed=2, random_diagonal=1, permute_nodes=1)
    assert t3.n_elems == 160 and t3.npe == 3
    assert sorted(np.unique(t3.conn)) == list(range(t3.n_nodes))
    mix = alls.generate(alls.GEN_MIXED, 40, 40, seed=2, random_diagonal=1, permute_nodes=1, split_fraction=0.5)
    nq = int((mix.elem_type == alls.Q4).sum()); nt = int((mix.elem_type == alls.T3).sum())
    assert nq + nt // 2 == 1600 and nt % 2 == 0 and 0.4 < nq / 1600 < 0.6
    assert (mix.conn.reshape(-1, 4)[mix.elem_type == alls.T3, 3] == -1).all()
    hx = alls.generate(alls.GEN_HEX20, 3, 2, 2, jitter=0.15, seed=2)
    # compact numbering: corners (n+1)^3-like plus mid-edge nodes, no orphans
    nc = 4 * 3 * 3
    ne = 3 * 3 * 3 + 4 * 2 * 3 + 4 * 3 * 2
    assert hx.n_nodes == nc + ne and hx.n_elems == 12
    assert sorted(np.unique(hx.conn)) == list(range(hx.n_nodes))


def test_jacobian_signs_of_benchmark_meshes():
    # Q4 with +-0.2h jitter cannot fold: at a corner the edge cross product is
    # >= (1-0.4)^2 - 0.4^2 > 0.  Triangles with +-0.3h jitter can invert when the
    # split diagonal is n1-n3; the reference's T3 path has no determinant clamp
    # (T3:184), so such elements simply contribute with negative area.
    from _checks import element_areas_2d
    q4 = alls.generate(alls.GEN_Q4, 60, 50, jitter=0.2, seed=5)
    assert element_areas_2d(q4).min() > 0
    t3 = alls.generate(alls.GEN_T3, 60, 50, jitter=0.2, seed=5, random_diagonal=1)
    assert element_areas_2d(t3).min() > 0
    t3b = alls.generate(alls.GEN_T3, 200, 200, jitter=0.3, seed=5, random_diagonal=1)
    frac = (element_areas_2d(t3b) <= 0).mean()
    assert frac < 1e-3


@pytest.mark.parametrize("kind,kw", [(alls.GEN_Q4, dict(nx=7, ny=5, E_sigma=0.5)),
                                     (alls.GEN_MIXED, dict(nx=6, ny=4, random_diagonal=1, E=3.0)),
                                     (alls.GEN_HEX20, dict(nx=2, ny=1, nz=3, jitter=0.1))])
def test_mesh_file_round_trip(tmp_path, kind, kw):
    m = alls.generate(kind, seed=4, **kw)
    path = tmp_path / "m.bin"
    alls.save_mesh(m, path)
    r = alls.load_mesh(path)
    for f in ("coords", "conn", "elem_type", "E_per_elem"):
        a, b = getattr(m, f), getattr(r, f)
        assert (a is None and b is None) or np.array_equal(a, b), f
    assert (r.dim, r.npe, r.uniform_type, r.E, r.nu, r.thickness) == (m.dim, m.npe, m.uniform_type, m.E, m.nu,
                                                                      m.thickness)


def test_mesh_file_rejects_garbage(tmp_path):
    path = tmp_path / "bad.bin"
    path.write_bytes(b"not a mesh file at all" * 8)
    with pytest.raises(alls.AllsError):
        alls.load_mesh(path)
    with pytest.raises(alls.AllsError):
        alls.load_mesh(tmp_path / "missing.bin")

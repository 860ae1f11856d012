"""GPU: the benchmark workloads at their full sizes, checked for correctness.

C2 (Q4 2000x2000) and C3 (T3 and mixed 1500x1500) are compared with the
serial oracle entry by entry (compare_csr: bit-identical row_ptr/col_ind,
values within 1e-12 * sum|contributions|), and the number of entries past
1e-12 relative is printed.  Hex20 80^3 (nnz 1,086,410,889) is too large for
the serial oracle (minutes of CPU), so both precisions are checked by
size-independent properties: the closed-form nnz, per-row sorted unique
columns, symmetry, K u = 0 for the six rigid-body modes, and the energy of
a uniform strain, sum_e V_e eps^T D eps with V_e the element volume under the
same 3x3x3 Gauss rule (exact: a linear field is reproduced by the element).
The FP32-MFMA result is also compared with the FP64 one (Frobenius <= 1e-6).
Meshes come from bench.py's own workload functions.
"""
import numpy as np
import pytest

import alls
import bench
import oracle
from _checks import (compare_csr, element_areas_2d, hex20_gauss_volumes, plane_stress_D, rigid_modes, solid_D,
                     uniform_strain_field)

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def invariants_2d(m, rp, ci, v):
    assert oracle.rows_sorted_unique_violations(rp, ci) == 0
    d, miss, vmax = oracle.symmetry(rp, ci, v)
    assert miss == 0 and d <= 1e-14 * vmax, (d, miss, vmax)
    for u in rigid_modes(m.coords, 2):
        assert np.abs(oracle.matvec(rp, ci, v, u)).max() < 1e-10 * vmax
    eps = np.array([[0.2, 0.15], [0.15, -0.1]])
    ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
    E = m.E_per_elem if m.E_per_elem is not None else np.full(m.n_elems, m.E)
    expect = (element_areas_2d(m) * E).sum() * m.thickness * (ev @ plane_stress_D(1.0, m.nu) @ ev)
    u = uniform_strain_field(m.coords, 2, eps)
    got = u @ oracle.matvec(rp, ci, v, u)
    assert abs(got - expect) < 1e-9 * abs(expect), (got, expect)


def test_c2_q4_2000_full_size(ctx):
    n = 2000
    m = bench.q4_workload(n)
    print("c2 generated", flush=True)
    out = ctx.assemble(m)
    gpu = out.download()
    out.free()
    rp, ci, v = gpu
    assert rp[-1] == 4 * (3 * n + 1) ** 2 == 144048004
    invariants_2d(m, rp, ci, v)
    print("c2 invariants ok", flush=True)
    met = compare_csr(gpu, oracle.assemble(m))
    print("c2 vs oracle", met, flush=True)


@pytest.mark.parametrize("name", ["t3", "mixed"])
def test_c3_1500_full_size_against_oracle(ctx, name):
    m = bench.c3_workload(name, 1500)
    print(f"c3 {name} generated: {m.n_elems} elements", flush=True)
    out = ctx.assemble(m)
    gpu = out.download()
    out.free()
    rp, ci, v = gpu
    assert oracle.rows_sorted_unique_violations(rp, ci) == 0
    d, miss, vmax = oracle.symmetry(rp, ci, v)
    assert miss == 0 and d <= 1e-14 * vmax, (d, miss, vmax)
    met = compare_csr(gpu, oracle.assemble(m))
    print(f"c3 {name} vs oracle", met, flush=True)


def test_c4_hex20_80_full_size_invariants(ctx):
    n = 80
    m = bench.hex20_workload(n)
    print("c4 generated", flush=True)
    eps = np.array([[0.3, 0.1, -0.2], [0.1, -0.4, 0.05], [-0.2, 0.05, 0.2]])
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
    vol = hex20_gauss_volumes(m.coords, m.conn)
    expect = vol.sum() * m.E * (ev @ solid_D(1.0, m.nu) @ ev)
    u_strain = uniform_strain_field(m.coords, 3, eps)
    modes = rigid_modes(m.coords, 3)
    v64 = None
    # (precision, symmetry bound, rigid-mode bound, energy bound), relative
    # to max |K|, max |K| and the energy: FP32 integration rounds each
    # Gauss-point product to 24 bits.
    for prec, sym_tol, rig_tol, en_tol in ((alls.PREC_F64, 1e-14, 1e-10, 1e-9),
                                           (alls.PREC_F32_MFMA, 1e-6, 1e-5, 1e-5)):
        out = ctx.assemble(m, precision=prec)
        rp, ci, v = out.download()
        out.free()
        print(f"c4 precision {prec} downloaded", flush=True)
        assert rp[-1] == 1086410889
        assert oracle.rows_sorted_unique_violations(rp, ci) == 0
        d, miss, vmax = oracle.symmetry(rp, ci, v)
        assert miss == 0 and d <= sym_tol * vmax, (prec, d, miss, vmax)
        rig = max(np.abs(oracle.matvec(rp, ci, v, u)).max() for u in modes)
        assert rig < rig_tol * vmax, (prec, rig, vmax)
        got = u_strain @ oracle.matvec(rp, ci, v, u_strain)
        assert abs(got - expect) < en_tol * abs(expect), (prec, got, expect)
        print(f"c4 precision {prec}: symmetry {d / vmax:.2e}, rigid {rig / vmax:.2e}, "
              f"energy rel {abs(got - expect) / abs(expect):.2e}", flush=True)
        if prec == alls.PREC_F64:
            v64 = v
        else:
            fro = oracle.frobenius_rel(v, v64)
            print(f"c4 FP32-MFMA vs FP64 Frobenius {fro:.3e}", flush=True)
            assert fro <= 1e-6
        del rp, ci, v


def test_hex20_reference_lattice_with_orphan_nodes(ctx):
    # The reference numbers every point of the (2nx+1)(2ny+1)(2nz+1) half-step
    # lattice (Hex:326-338); face and body centres belong to no element, so
    # their DOF rows are empty.
    nx, ny, nz = 3, 2, 2
    LX, LY, LZ = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    k, j, i = np.meshgrid(np.arange(LZ), np.arange(LY), np.arange(LX), indexing="ij")
    pts = np.column_stack([0.5 * i.ravel() / nx, 0.5 * j.ravel() / ny, 0.5 * k.ravel() / nz])
    rng = np.random.default_rng(12)
    inner = ((i > 0) & (i < LX - 1) & (j > 0) & (j < LY - 1) & (k > 0) & (k < LZ - 1)).ravel()
    pts[inner] += rng.uniform(-0.04, 0.04, (inner.sum(), 3))
    off = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2), (1, 0, 0),
           (2, 1, 0), (1, 2, 0), (0, 1, 0), (1, 0, 2), (2, 1, 2), (1, 2, 2), (0, 1, 2), (0, 0, 1), (2, 0, 1),
           (2, 2, 1), (0, 2, 1)]
    conn = np.array([[(2 * c + dz) * LX * LY + (2 * b + dy) * LX + (2 * a + dx) for dx, dy, dz in off]
                     for c in range(nz) for b in range(ny) for a in range(nx)], np.int32)
    m = alls.MeshArrays(pts.ravel(), conn, 3, 20, uniform_type=alls.HEX20, E=1.0, nu=0.3)
    gpu = ctx.assemble(m).download()
    used = np.zeros(m.n_nodes, bool)
    used[conn.ravel()] = True
    assert (~used).sum() == nx * ny * nz + nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz
    lens = np.diff(gpu[0]).reshape(-1, 3)
    assert (lens[~used] == 0).all() and (lens[used] > 0).all()
    compare_csr(gpu, oracle.assemble(m))

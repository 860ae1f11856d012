"""Pins the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests, fixtures or golden vectors for this path
(it only prints a host-vs-GPU comparison at run time, Acc:1130-1186), and
its code cannot be compiled here (CUDA headers and nvcc are absent), so
each oracle function is pinned by mathematics chosen so that a dropped
term, a wrong sign/index or a transposed operand fails a check.
"""
import json
import os

import numpy as np
import pytest

import alls
import oracle
from _checks import (HEX_NAT, csr_matvec, csr_symmetry_error, element_areas_2d, hex20_dshape_numpy,
                     plane_stress_D, rigid_modes, solid_D, uniform_strain_field)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ Q4 ---
def top88_ke(nu):
    """Closed-form unit-square Q4 stiffness (Andreassen et al. 2011, E=1,t=1)."""
    k = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8,
                  -1 / 4 + nu / 12, -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1],
           [3, 6, 5, 0, 7, 2, 1, 4], [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6],
           [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    return k[np.array(idx)] / (1 - nu * nu)


@pytest.mark.parametrize("nu", [0.0, 0.25, 0.3, 0.45])
def test_q4_unit_square_top88(nu):
    ke, _ = oracle.ke_q4([0, 1, 1, 0], [0, 0, 1, 1], E=1.0, nu=nu, t=1.0)
    assert np.abs(ke - top88_ke(nu)).max() < 1e-15


def test_q4_scaling_E_thickness_and_size():
    x, y = np.array([0, 1, 1, 0.0]), np.array([0, 0, 1, 1.0])
    k1, _ = oracle.ke_q4(x, y, 1.0, 0.3, 1.0)
    k2, _ = oracle.ke_q4(3 * x + 5, 3 * y - 2, 2.5, 0.3, 0.1)   # plane stress: size-invariant
    assert np.allclose(k2, 2.5 * 0.1 * k1, rtol=1e-13, atol=1e-15)


def q4_ke_numpy(x, y, E, nu, t, ngp):
    """Independent Q4 stiffness: B-matrix form with an n-point Gauss rule."""
    g, w = np.polynomial.legendre.leggauss(ngp)
    D = plane_stress_D(E, nu)
    K = np.zeros((8, 8))
    for xi, wi in zip(g, w):
        for eta, wj in zip(g, w):
            dN = 0.25 * np.array([[-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)],
                                  [-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)]])
            J = dN @ np.column_stack([x, y])
            G = np.linalg.solve(J, dN)
            B = np.zeros((3, 8))
            B[0, 0::2] = G[0]; B[1, 1::2] = G[1]; B[2, 0::2] = G[1]; B[2, 1::2] = G[0]
            K += B.T @ D @ B * np.linalg.det(J) * wi * wj * t
    return K


def test_q4_parallelogram_against_independent_rule():
    # Affine element: integrand is polynomial, so 2x2 and 4x4 Gauss agree.
    x = np.array([0.1, 1.3, 1.7, 0.5]); y = np.array([-0.2, 0.1, 1.0, 0.7])
    ke, ka = oracle.ke_q4(x, y, 2.0, 0.27, 0.5)
    ref = q4_ke_numpy(x, y, 2.0, 0.27, 0.5, 4)
    assert np.abs(ke - ref).max() < 1e-13 * np.abs(ref).max()
    assert (ka >= np.abs(ke) - 1e-15).all()


def test_q4_jittered_null_space_and_symmetry():
    rng = np.random.default_rng(3)
    x = np.array([0, 1, 1, 0.0]) + rng.uniform(-0.2, 0.2, 4)
    y = np.array([0, 0, 1, 1.0]) + rng.uniform(-0.2, 0.2, 4)
    ke, _ = oracle.ke_q4(x, y, 1.0, 0.3, 1.0)
    assert np.abs(ke - ke.T).max() < 1e-15
    for u in rigid_modes(np.column_stack([x, y]).reshape(-1), 2):
        assert np.abs(ke @ u).max() < 1e-14
    assert np.linalg.matrix_rank(ke, tol=1e-10) == 5
    # independent rule on a general quad agrees to quadrature accuracy only
    assert np.abs(ke - q4_ke_numpy(x, y, 1.0, 0.3, 1.0, 2)).max() < 1e-14


# ------------------------------------------------------------------ T3 ---
def cst_closed_form(x, y, E, nu, t):
    """t * A * B^T D B with B from the edge vectors (not the Jacobian path)."""
    x1, x2, x3 = x; y1, y2, y3 = y
    A2 = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    b = np.array([y2 - y3, y3 - y1, y1 - y2]); c = np.array([x3 - x2, x1 - x3, x2 - x1])
    B = np.zeros((3, 6))
    B[0, 0::2] = b; B[1, 1::2] = c; B[2, 0::2] = c; B[2, 1::2] = b
    B /= A2
    return t * 0.5 * A2 * B.T @ plane_stress_D(E, nu) @ B


@pytest.mark.parametrize("seed", range(5))
def test_t3_cst_closed_form(seed):
    rng = np.random.default_rng(seed)
    x = np.array([0, 1, 0.2]) + rng.uniform(-0.2, 0.2, 3)
    y = np.array([0, 0.1, 1.0]) + rng.uniform(-0.2, 0.2, 3)
    ke, ka = oracle.ke_t3(x, y, 1.7, 0.31, 0.2)
    ref = cst_closed_form(x, y, 1.7, 0.31, 0.2)
    assert np.abs(ke - ref).max() < 1e-14 * np.abs(ref).max()
    assert np.linalg.matrix_rank(ke, tol=1e-10) == 3
    for u in rigid_modes(np.column_stack([x, y]).reshape(-1), 2):
        assert np.abs(ke @ u).max() < 1e-14


# --------------------------------------------------------------- Hex20 ---
def hex20_ke_numpy(X, E, nu, ngp):
    g, w = np.polynomial.legendre.leggauss(ngp)
    D = solid_D(E, nu)
    K = np.zeros((60, 60))
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for c, wc in zip(g, w):
                dN = hex20_dshape_numpy((a, b, c))
                J = dN @ X
                G = np.linalg.solve(J, dN)
                B = np.zeros((6, 60))
                B[0, 0::3] = G[0]; B[1, 1::3] = G[1]; B[2, 2::3] = G[2]
                B[3, 0::3] = G[1]; B[3, 1::3] = G[0]
                B[4, 1::3] = G[2]; B[4, 2::3] = G[1]
                B[5, 0::3] = G[2]; B[5, 2::3] = G[0]
                K += B.T @ D @ B * np.linalg.det(J) * wa * wb * wc
    return K


def brick(affine=True, seed=0):
    X = (HEX_NAT + 1) / 2
    A = np.array([[1.2, 0.1, 0.05], [0.0, 0.9, -0.1], [0.1, 0.2, 1.1]])
    X = X @ A.T + np.array([0.3, -0.1, 0.2])
    if not affine:
        X = X + np.random.default_rng(seed).uniform(-0.04, 0.04, X.shape)
    return X


def test_hex20_affine_against_independent_4x4x4():
    X = brick(affine=True)
    ke, ka = oracle.ke_hex20(X[:, 0], X[:, 1], X[:, 2], 1.3, 0.3)
    ref = hex20_ke_numpy(X, 1.3, 0.3, 4)
    # Gauss literals are the reference's 15-digit ones: ~1e-15 relative.
    assert np.abs(ke - ref).max() < 2e-14 * np.abs(ref).max()
    assert (ka >= np.abs(ke) * (1 - 1e-12)).all()


@pytest.mark.parametrize("affine", [True, False])
def test_hex20_invariants(affine):
    X = brick(affine=affine, seed=5)
    ke, _ = oracle.ke_hex20(X[:, 0], X[:, 1], X[:, 2], 1.0, 0.3)
    s = np.abs(ke).max()
    assert np.abs(ke - ke.T).max() < 1e-14 * s
    for u in rigid_modes(X.reshape(-1), 3):
        assert np.abs(ke @ u).max() < 1e-13 * s
    assert np.linalg.matrix_rank(ke, tol=1e-9 * s) == 54
    # Uniform strain: u = eps x reproduced exactly; energy = V eps^T D eps.
    eps = np.array([[0.3, 0.1, -0.2], [0.1, -0.4, 0.05], [-0.2, 0.05, 0.2]])
    u = uniform_strain_field(X.reshape(-1), 3, eps)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
    g, w = np.polynomial.legendre.leggauss(3)
    V = sum(np.linalg.det(hex20_dshape_numpy((a, b, c)) @ X) * wa * wb * wc
            for a, wa in zip(g, w) for b, wb in zip(g, w) for c, wc in zip(g, w))
    assert abs(u @ ke @ u - V * ev @ solid_D(1.0, 0.3) @ ev) < 1e-13 * abs(V * ev @ solid_D(1.0, 0.3) @ ev)


def test_hex20_f32_restatement_close_to_f64():
    X = brick(affine=False, seed=2)
    k64, _ = oracle.ke_hex20(X[:, 0], X[:, 1], X[:, 2], 1.0, 0.3)
    k32 = oracle.ke_hex20_f32(X[:, 0], X[:, 1], X[:, 2], 1.0, 0.3)
    fro = np.linalg.norm(k32 - k64) / np.linalg.norm(k64)
    assert 1e-9 < fro < 1e-5


# ------------------------------------------------------------- pattern ---
MESHES = [
    (alls.GEN_Q4, dict(nx=4, ny=3)),
    (alls.GEN_T3, dict(nx=4, ny=3, random_diagonal=1, permute_nodes=1)),
    (alls.GEN_MIXED, dict(nx=5, ny=4, random_diagonal=1, permute_nodes=1)),
    (alls.GEN_HEX20, dict(nx=2, ny=3, nz=2)),
]


@pytest.mark.parametrize("kind,kw", MESHES)
def test_pattern_matches_bruteforce(kind, kw):
    m = alls.generate(kind, jitter=0.2, seed=11, **kw)
    rp, ci = oracle.pattern(m)
    rb, cb = oracle.pattern_bruteforce(m)
    assert np.array_equal(rp, rb) and np.array_equal(ci, cb)
    for r in range(rp.size - 1):
        row = ci[rp[r]:rp[r + 1]]
        assert (np.diff(row) > 0).all()


@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_pattern_closed_form_nnz(n):
    q4 = alls.generate(alls.GEN_Q4, n, n)
    assert oracle.pattern(q4)[0][-1] == 4 * (3 * n + 1) ** 2
    t3 = alls.generate(alls.GEN_T3, n, n)             # reference diagonal n0-n2
    nn = (n + 1) ** 2
    assert oracle.pattern(t3)[0][-1] == 4 * (nn + 2 * (2 * n * (n + 1) + n * n))


# ------------------------------------------------------------ assembly ---
@pytest.mark.parametrize("kind,kw", MESHES)
def test_assembly_equals_dense_sum(kind, kw):
    m = alls.generate(kind, jitter=0.2, seed=5, E_sigma=0.5, **kw)
    rp, ci, v, av = oracle.assemble(m)
    K = oracle.assemble_dense(m)
    Kc = np.zeros_like(K)
    for r in range(rp.size - 1):
        Kc[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    assert np.array_equal(K, Kc)
    assert (av >= np.abs(v) * (1 - 1e-12)).all()


@pytest.mark.parametrize("kind,kw", MESHES)
def test_assembly_global_invariants(kind, kw):
    m = alls.generate(kind, jitter=0.2, seed=9, **kw)
    rp, ci, v, _ = oracle.assemble(m)
    s = np.abs(v).max()
    assert csr_symmetry_error(rp, ci, v) < 1e-14
    for u in rigid_modes(m.coords, m.dim):
        assert np.abs(csr_matvec(rp, ci, v, u)).max() < 1e-12 * s
    if m.dim == 2:
        eps = np.array([[0.2, 0.15], [0.15, -0.1]])
        ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
        expect = element_areas_2d(m).sum() * m.thickness * ev @ plane_stress_D(m.E, m.nu) @ ev
    else:
        eps = np.array([[0.3, 0.1, -0.2], [0.1, -0.4, 0.05], [-0.2, 0.05, 0.2]])
        ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
        expect = 1.0 * ev @ solid_D(m.E, m.nu) @ ev   # unit cube
    u = uniform_strain_field(m.coords, m.dim, eps)
    got = u @ csr_matvec(rp, ci, v, u)
    assert abs(got - expect) < 1e-11 * abs(expect)


def test_assembly_planted_error_is_caught():
    # A transposed block in one element must break the symmetry/energy pins.
    m = alls.generate(alls.GEN_Q4, 4, 4, jitter=0.2, seed=1)
    rp, ci, v, _ = oracle.assemble(m)
    v2 = v.copy()
    v2[5] *= 1.0 + 1e-9
    assert csr_symmetry_error(rp, ci, v2) > 1e-12


def test_precision_metrics_hand_computed():
    ref = np.array([2.0, -4.0, 0.0, 1.0])
    test = np.array([2.0, -3.0, 5.0, 1.5])
    p = oracle.precision(ref, test)
    assert p["max_abs"] == 1.0 and p["max_rel"] == 0.5
    assert p["mean_abs"] == pytest.approx(1.5 / 4) and p["mean_rel"] == pytest.approx((0.25 + 0.5) / 4)
    assert p["frobenius_rel"] == pytest.approx(np.sqrt(1 + 0.25) / np.sqrt(21))


@pytest.mark.parametrize("kind,kw", MESHES)
def test_greedy_coloring_valid(kind, kw):
    m = alls.generate(kind, **kw)
    n, col = oracle.color_greedy(m)
    assert oracle.coloring_conflicts(m, col) == 0
    assert col.max() + 1 == n
    assert oracle.coloring_conflicts(m, m.structured_color) == 0
    bad = col.copy()
    bad[1] = bad[0]                      # elements 0 and 1 share nodes in every kind
    assert oracle.coloring_conflicts(m, bad) > 0


def test_golden_c1_fixture():
    """C1 (Q4 64x64, E=1, nu=0.3, t=1) against closed-form top88 expectations
    written by tests/golden/make_golden.py (independent of the oracle)."""
    with open(os.path.join(GOLDEN, "c1_q4_64.json")) as f:
        g = json.load(f)
    m = alls.generate(alls.GEN_Q4, 64, 64, E=1.0, nu=0.3, thickness=1.0)
    rp, ci, v, _ = oracle.assemble(m)
    assert int(rp[-1]) == g["nnz"] == 148996
    assert int(ci.astype(np.int64).sum()) == g["col_ind_sum"]
    assert abs(float(v.sum()) - g["val_sum"]) <= 1e-12 * g["val_abs_sum"]
    assert abs(float(np.abs(v).sum()) - g["val_abs_sum"]) <= 1e-12 * g["val_abs_sum"]
    assert np.allclose(v[:len(g["val_head"])], g["val_head"], rtol=1e-14, atol=1e-16)
    r = 2 * (32 * 65 + 32)
    diag = v[rp[r]:rp[r + 1]][ci[rp[r]:rp[r + 1]] == r][0]
    assert abs(diag - g["interior_diag"]) < 1e-14


def test_csr_checks_catch_planted_errors():
    # The full-size GPU tests rest on these checks: each must see a planted
    # mistake of the kind it guards against.
    m = alls.generate(alls.GEN_T3, 12, 9, jitter=0.2, seed=4, random_diagonal=1, permute_nodes=1, E_sigma=0.5)
    rp, ci, v, _ = oracle.assemble(m)
    x = np.random.default_rng(1).standard_normal(rp.size - 1)
    assert np.allclose(oracle.matvec(rp, ci, v, x), csr_matvec(rp, ci, v, x), rtol=0, atol=1e-12)
    assert oracle.rows_sorted_unique_violations(rp, ci) == 0
    d, miss, vmax = oracle.symmetry(rp, ci, v)
    assert miss == 0 and d <= 1e-14 * vmax
    assert abs(d / vmax - csr_symmetry_error(rp, ci, v)) < 1e-15
    bad = ci.copy()
    bad[rp[7] + 1], bad[rp[7] + 2] = bad[rp[7] + 2], bad[rp[7] + 1]   # two columns swapped
    assert oracle.rows_sorted_unique_violations(rp, bad) == 1
    bad = ci.copy()
    bad[rp[7] + 1] = bad[rp[7]]                                      # a repeated column
    assert oracle.rows_sorted_unique_violations(rp, bad) == 1
    w = v.copy()
    k = next(k for k in range(rp[30], rp[31]) if ci[k] != 30)
    w[k] += 1e-9 * vmax                                              # one off-diagonal value off
    assert oracle.symmetry(rp, ci, w)[0] > 0.5e-9 * vmax
    assert oracle.frobenius_rel(w, v) > 0 and oracle.frobenius_rel(v, v) == 0


def test_hex20_gauss_volumes_tile_the_cube():
    from _checks import hex20_gauss_volumes
    m = alls.generate(alls.GEN_HEX20, 3, 2, 2, jitter=0.15, seed=7)
    vol = hex20_gauss_volumes(m.coords, m.conn)
    assert (vol > 0).all() and abs(vol.sum() - 1.0) < 1e-13
    X = m.coords.reshape(-1, 3)
    c = m.conn.reshape(-1, 20)
    for e in (0, 5):
        g, w = np.polynomial.legendre.leggauss(3)
        ref = sum(np.linalg.det(hex20_dshape_numpy((a, b, cc)) @ X[c[e]]) * wa * wb * wc
                  for a, wa in zip(g, w) for b, wb in zip(g, w) for cc, wc in zip(g, w))
        assert abs(vol[e] - ref) < 1e-15

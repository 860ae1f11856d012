"""GPU: cached pattern + value-only re-assembly, GPU colouring, coloured path.

All values are checked against the oracle's serial assembly with the same
tolerance as the A-LLS path (_checks.VAL_RTOL, 1e-12 * sum|contributions|).
"""
import numpy as np
import pytest

import alls
import oracle
from _checks import compare_csr

pytestmark = pytest.mark.gpu

MESHES = [
    ("q4", alls.GEN_Q4, dict(nx=70, ny=50, jitter=0.2, seed=31, E_sigma=0.5)),
    ("t3", alls.GEN_T3, dict(nx=50, ny=40, jitter=0.3, seed=32, random_diagonal=1, permute_nodes=1)),
    ("mixed", alls.GEN_MIXED, dict(nx=50, ny=40, jitter=0.3, seed=33, random_diagonal=1, permute_nodes=1,
                                   E_sigma=0.5)),
    ("hex20", alls.GEN_HEX20, dict(nx=4, ny=4, nz=3, jitter=0.15, seed=34, E_sigma=0.5)),
]


@pytest.mark.parametrize("name,kind,kw", MESHES, ids=[m[0] for m in MESHES])
def test_value_only_reassembly_with_new_E(ctx, name, kind, kw):
    m = alls.generate(kind, **kw)
    pat = alls.Pattern(ctx, m)
    compare_csr(pat.out.download(), oracle.assemble(m))
    rng = np.random.default_rng(5)
    for _ in range(2):
        E2 = np.exp(0.5 * rng.standard_normal(m.n_elems))
        out = pat.values(E2)
        got = out.download()
        compare_csr(got, oracle.assemble(m.with_E(E2)))
        out.free()
        # Same blocks, same element-order sums: bitwise the plain assembly.
        plain = ctx.assemble(m.with_E(E2))
        for a, b in zip(got, plain.download()):
            assert np.array_equal(a, b)
        plain.free()
    pat.free()


@pytest.mark.parametrize("name,kind,kw", MESHES, ids=[m[0] for m in MESHES])
def test_gpu_coloring_valid_and_colored_assembly(ctx, name, kind, kw):
    m = alls.generate(kind, **kw)
    pat = alls.Pattern(ctx, m)
    col = pat.color_gpu()
    c = col.download(m.n_elems)
    assert oracle.coloring_conflicts(m, c) == 0
    assert col.n_colors == c.max() + 1
    greedy, _ = oracle.color_greedy(m)
    assert col.n_colors <= 2 * greedy + 2
    print(name, "JPL colours", col.n_colors, "greedy", greedy)
    ref = oracle.assemble(m)
    compare_csr(pat.colored(col).download(), ref)
    # the structured cell-parity colouring of the generator
    sc = pat.color_from_host(m.structured_color)
    assert sc.n_colors == m.n_structured_colors
    compare_csr(pat.colored(sc).download(), ref)


def test_coloring_with_conflict_is_rejected(ctx):
    m = alls.generate(alls.GEN_Q4, 10, 10)
    pat = alls.Pattern(ctx, m)
    bad = m.structured_color.copy()
    bad[1] = bad[0]
    with pytest.raises(alls.AllsError) as ei:
        pat.color_from_host(bad)
    assert ei.value.status == alls.E_INVALID


def test_value_only_output_reuse_and_pattern_switch(ctx):
    a = alls.generate(alls.GEN_Q4, 30, 20, jitter=0.2, seed=1, E_sigma=0.5)
    b = alls.generate(alls.GEN_T3, 30, 20, jitter=0.2, seed=2, random_diagonal=1, E_sigma=0.5)
    pa, pb = alls.Pattern(ctx, a), alls.Pattern(ctx, b)
    out = pa.values(a.E_per_elem)
    out = pb.values(b.E_per_elem, out=out)      # same buffers, other pattern
    compare_csr(out.download(), oracle.assemble(b))
    out = pa.values(a.E_per_elem, out=out)
    compare_csr(out.download(), oracle.assemble(a))


@pytest.mark.parametrize("prec", [alls.PREC_F64, alls.PREC_F32_MFMA], ids=["f64", "f32"])
def test_hex20_value_only_rows_with_many_elements(ctx, prec):
    """Every element of a 2x2x2 brick listed twice: the centre vertex row
    has 16 incident elements, more than the value-only gather takes in one
    round of loads (9), so it goes through the three-at-a-time loop; the
    edge and face rows (8 or fewer) through the one-round path.  Both equal
    the plain assembly bitwise, and the oracle."""
    m0 = alls.generate(alls.GEN_HEX20, 2, 2, 2, jitter=0.1, seed=11)
    conn = m0.conn.reshape(-1, 20)
    conn2 = np.concatenate([conn, conn[::-1]])
    E = np.exp(0.5 * np.random.default_rng(5).standard_normal(conn2.shape[0]))
    m = alls.MeshArrays(m0.coords, conn2.reshape(-1), 3, 20, m0.uniform_type, None, 1.0, E, m0.nu,
                        m0.thickness)
    inc = np.bincount(conn2.reshape(-1), minlength=m.n_nodes)
    assert inc.max() == 16 and ((inc > 0) & (inc <= 9)).any()
    plain = ctx.assemble(m, precision=prec)
    got = plain.download()
    plain.free()
    ref = oracle.assemble(m)
    if prec == alls.PREC_F64:
        compare_csr(got, ref)
    else:
        assert np.linalg.norm(got[2] - ref[2]) / np.linalg.norm(ref[2]) < 1e-6
    pat = alls.Pattern(ctx, m, precision=prec)
    out = pat.values(E)
    for a, b in zip(out.download(), got):
        assert np.array_equal(a, b)
    out.free()
    pat.free()


def test_hex20_f32_value_only(ctx):
    m = alls.generate(alls.GEN_HEX20, 4, 3, 3, jitter=0.15, seed=7)
    pat = alls.Pattern(ctx, m, precision=alls.PREC_F32_MFMA)
    E2 = np.exp(0.5 * np.random.default_rng(1).standard_normal(m.n_elems))
    v = pat.values(E2).download()[2]
    ref = oracle.assemble(m.with_E(E2))[2]
    assert np.linalg.norm(v - ref) / np.linalg.norm(ref) < 1e-6
    assert np.array_equal(v, ctx.assemble(m.with_E(E2), precision=alls.PREC_F32_MFMA).download()[2])

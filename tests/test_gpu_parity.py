"""GPU parity: the HIP A-LLS path through the C ABI against the CPU oracle.

Bar (see _checks.py): row_ptr/col_ind bit-identical to the oracle's serial
restatement of the reference (Acc:329-380 with the std::set pattern of
Acc:605-645 / Hex:364-392); values within 1e-12 * sum|contributions| per
entry; FP32-MFMA Hex20 within 1e-6 relative Frobenius of the FP64 oracle.
"""
import numpy as np
import pytest

import alls
import oracle
from _checks import (compare_csr, csr_matvec, csr_symmetry_error, element_areas_2d, plane_stress_D,
                     rigid_modes, solid_D, uniform_strain_field)

pytestmark = pytest.mark.gpu

CASES = [
    ("c1_q4_64", alls.GEN_Q4, dict(nx=64, ny=64)),
    ("q4_jitter_lognormal", alls.GEN_Q4, dict(nx=90, ny=70, jitter=0.2, seed=11, E_sigma=0.5)),
    ("t3_randdiag_perm", alls.GEN_T3, dict(nx=60, ny=50, jitter=0.3, seed=12, random_diagonal=1,
                                          permute_nodes=1)),
    ("t3_reference_diag", alls.GEN_T3, dict(nx=40, ny=30)),
    ("mixed_perm", alls.GEN_MIXED, dict(nx=60, ny=50, jitter=0.3, seed=13, random_diagonal=1,
                                        permute_nodes=1, E_sigma=0.5)),
    ("hex20_small", alls.GEN_HEX20, dict(nx=5, ny=4, nz=3, jitter=0.15, seed=14, E_sigma=0.5)),
    ("hex20_single", alls.GEN_HEX20, dict(nx=1, ny=1, nz=1, jitter=0.0, seed=0)),
    ("q4_single", alls.GEN_Q4, dict(nx=1, ny=1)),
]


def run_case(ctx, kind, kw, **opts):
    m = alls.generate(kind, **kw)
    out = ctx.assemble(m, **opts)
    gpu = out.download()
    out.free()
    return m, gpu


@pytest.mark.parametrize("name,kind,kw", CASES, ids=[c[0] for c in CASES])
def test_parity_against_oracle(ctx, name, kind, kw):
    print(f"case {name}", flush=True)
    m, gpu = run_case(ctx, kind, kw)
    ref = oracle.assemble(m)
    met = compare_csr(gpu, ref)
    print(name, met)


def test_hex20_f32_mfma_frobenius(ctx):
    m, gpu = run_case(ctx, alls.GEN_HEX20, dict(nx=6, ny=5, nz=4, jitter=0.15, seed=21, E_sigma=0.5),
                      precision=alls.PREC_F32_MFMA)
    rp, ci, v, _ = oracle.assemble(m)
    assert np.array_equal(gpu[0], rp) and np.array_equal(gpu[1], ci)
    fro = np.linalg.norm(gpu[2] - v) / np.linalg.norm(v)
    print("f32 frobenius", fro)
    assert fro <= 1e-6


REPEATED = [
    (alls.GEN_Q4, dict(nx=7, ny=6, jitter=0.1, seed=3)),
    (alls.GEN_T3, dict(nx=7, ny=6, jitter=0.1, seed=3, random_diagonal=1)),
    (alls.GEN_MIXED, dict(nx=7, ny=6, jitter=0.1, seed=3, random_diagonal=1)),
    (alls.GEN_HEX20, dict(nx=3, ny=2, nz=2, jitter=0.1, seed=3)),
]


def repeated_node_mesh(kind, kw):
    """The generated mesh with one element's second node replaced by its
    first.  In a mixed mesh that element is a quadrilateral: a triangle that
    lists a node twice has zero area, and the reference's CST formula
    (T3:177-205, no determinant clamp) gives NaN blocks for it, which every
    path reproduces (checked by position only)."""
    m = alls.generate(kind, **kw)
    conn = m.conn.reshape(m.n_elems, m.npe).copy()
    e = m.n_elems // 2
    if m.elem_type is not None:
        e += int(np.argmax(m.elem_type[e:] == alls.Q4))
    conn[e, 1] = conn[e, 0]
    return e, alls.MeshArrays(m.coords, conn, m.dim, m.npe, m.uniform_type, m.elem_type, m.E, m.E_per_elem, m.nu,
                              m.thickness)


@pytest.mark.parametrize("kind,kw", REPEATED, ids=["q4", "t3", "mixed", "hex20"])
def test_element_with_a_repeated_node(ctx, kind, kw):
    # A degenerate element (one node id twice) puts two contributions of one
    # element on the same (row, column): the list layouts must keep both (the
    # Hex20 group merge falls back to scanning its groups for such a row),
    # and the two are summed in the element's (row node, column node) order,
    # the serial order, so repeated runs give the same bits.
    e, m = repeated_node_mesh(kind, kw)
    print(f"repeated node kind {kind} element {e}", flush=True)
    out = ctx.assemble(m)
    gpu = out.download()
    compare_csr(gpu, oracle.assemble(m))
    for _ in range(2):
        out = ctx.assemble(m, out=out)
        assert same_csr(out.download(), gpu)
    out.free()


def same_csr(a, b):
    """Bitwise equal CSR triples (NaN values compared by position)."""
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            and np.array_equal(a[2], b[2], equal_nan=True))


def repeated_node_fan():
    """40 elements around node 0 (a 121-entry row, past the register path):
    triangles, except element 3, the quadrilateral (0, 4, 5, 5) that lists
    node 5 twice.  Row 0 then holds two blocks of one element on column 5,
    and row 5 sees the element twice in its incidence list."""
    k = 40
    m = fan_mesh(k)
    conn = np.full((k, 4), -1, np.int32)
    conn[:, :3] = m.conn.reshape(k, 3)
    conn[3] = [0, 4, 5, 5]
    types = np.full(k, alls.T3, np.uint8)
    types[3] = alls.Q4
    return alls.MeshArrays(m.coords, conn, 2, 4, uniform_type=alls.Q4, elem_type=types)


@pytest.mark.parametrize("case", ["q4", "t3", "mixed", "hex20", "fan"])
def test_value_only_and_cache_hit_with_a_repeated_node(ctx, case):
    # The cached-pattern path (and a server cache hit, which runs it) must
    # give the plain assembly's bits for a mesh with a degenerate element.
    if case == "fan":
        m = repeated_node_fan()
    else:
        kind, kw = REPEATED[["q4", "t3", "mixed", "hex20"].index(case)]
        m = repeated_node_mesh(kind, kw)[1]
    print(f"repeated node value-only case {case}", flush=True)
    E2 = np.exp(0.5 * np.random.default_rng(21).standard_normal(m.n_elems))
    m2 = m.with_E(E2)
    out = ctx.assemble(m2)
    plain = out.download()
    out.free()
    compare_csr(plain, oracle.assemble(m2))
    pat = alls.Pattern(ctx, m)
    pat.out.free()
    out = pat.values(E2)
    assert same_csr(out.download(), plain)
    out.free()
    pat.free()
    with alls.Server(n_streams=2, cache_policy=alls.CACHE_EAGER) as srv:
        srv.wait(srv.submit(m, key=3), keep_output=False)
        csr, st = srv.wait(srv.submit(m2, key=3))
        assert st["cache_hit"] == 1
        got = srv.download(csr)
        srv.release(csr)
    assert same_csr(got, plain)


def test_hex20_invalid_node_with_overflow_chunks_fails_fast(ctx):
    # Elements with a bad node reserve list slots before the check (FP64);
    # with 40-slot chunks (vertex rows of 160 entries take 3 overflow chunks)
    # some of those reservations own a chunk, which they must still publish,
    # or the other elements of that row would wait for it until their spin
    # limit.  The call reports the bad node at once and the context stays
    # usable.
    import time
    m = alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=8)
    compare_csr(ctx.assemble(m, chunk_slots=40).download(), oracle.assemble(m))
    conn = m.conn.reshape(m.n_elems, 20).copy()
    conn[::3, 19] = m.n_nodes + 5
    bad = alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, None, m.nu, m.thickness)
    t0 = time.perf_counter()
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble(bad, chunk_slots=40)
    dt = time.perf_counter() - t0
    assert ei.value.status == alls.E_INVALID
    assert dt < 0.1, dt
    compare_csr(ctx.assemble(m, chunk_slots=40).download(), oracle.assemble(m))


def test_hex20_rows_beyond_the_short_count_table(ctx):
    # Every element listed twice: vertex rows get 320 entries, more than the
    # first count launch's table holds, so they are counted by the second
    # launch and merged by the LDS kernel.
    m = alls.generate(alls.GEN_HEX20, 2, 2, 2, jitter=0.1, seed=9)
    conn = np.concatenate([m.conn.reshape(m.n_elems, 20)] * 2)
    m = alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, None, m.nu, m.thickness)
    out = ctx.assemble(m)
    assert ctx.stats()["big_rows"] >= 1
    compare_csr(out.download(), oracle.assemble(m))


def test_hex20_rows_of_four_register_slots(ctx):
    # Two of the eight elements around the centre vertex of a 2x2x2 brick
    # listed twice (as new elements): that row gets 200 entries, so the
    # segment merge runs with four 60-entry register slots.
    m = alls.generate(alls.GEN_HEX20, 2, 2, 2, jitter=0.1, seed=11, E_sigma=0.5)
    c = m.conn.reshape(m.n_elems, 20)
    conn = np.concatenate([c, c[[0, 5]]])
    E = None if m.E_per_elem is None else np.concatenate([m.E_per_elem, m.E_per_elem[[0, 5]]])
    m = alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, E, m.nu, m.thickness)
    out = ctx.assemble(m)
    compare_csr(out.download(), oracle.assemble(m))


def test_hex20_row_with_more_columns_than_the_segment_merge_holds(ctx):
    # Six spread-out elements away from a corner node take it in place of
    # one of their own nodes: its row gets 113 columns (over 96), so the
    # segment merge leaves it to the generic kernel, and the rows it was
    # added to keep their shape.
    m = alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=12)
    conn = m.conn.reshape(m.n_elems, 20).copy()
    x = int(conn[0, 0])
    far = [e for e in range(m.n_elems) if x not in conn[e]][2::4][:6]
    conn[far, 7] = x
    m = alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, None, m.nu, m.thickness)
    out = ctx.assemble(m)
    rp, ci, _ = out.download()
    assert (rp[3 * x + 1] - rp[3 * x]) // 3 > 96
    compare_csr((rp, ci, _), oracle.assemble(m))


@pytest.mark.parametrize("prec", [alls.PREC_F64, alls.PREC_F32_MFMA], ids=["f64", "f32"])
def test_hex20_odd_chunk_size(ctx, prec):
    # 161-slot chunks: half of the rows start their groups at odd list
    # indices (the FP64 lane-pair stores fall back to single stores there)
    # and a group at slot 160 spans two chunks.
    m = alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=5, E_sigma=0.5)
    out = ctx.assemble(m, chunk_slots=161, precision=prec)
    assert ctx.stats()["chunk_slots"] == 161
    gpu = out.download()
    out.free()
    if prec == alls.PREC_F64:
        compare_csr(gpu, oracle.assemble(m))
    else:
        rp, ci, v, _ = oracle.assemble(m)
        assert np.array_equal(gpu[0], rp) and np.array_equal(gpu[1], ci)
        assert np.linalg.norm(gpu[2] - v) / np.linalg.norm(v) <= 1e-6


def test_repeat_runs_are_bitwise_identical(ctx):
    m = alls.generate(alls.GEN_T3, 80, 80, jitter=0.3, seed=3, random_diagonal=1, permute_nodes=1)
    outs = []
    for _ in range(3):
        o = ctx.assemble(m)
        outs.append(o.download())
        o.free()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_shuffled_element_order_same_csr(ctx):
    m = alls.generate(alls.GEN_MIXED, 50, 40, jitter=0.2, seed=8, random_diagonal=1, E_sigma=0.5)
    perm = np.random.default_rng(0).permutation(m.n_elems)
    conn = m.conn.reshape(-1, 4)[perm]
    m2 = alls.MeshArrays(m.coords, conn, 2, 4, elem_type=m.elem_type[perm], E_per_elem=m.E_per_elem[perm],
                         nu=m.nu, thickness=m.thickness)
    a = ctx.assemble(m).download()
    b = ctx.assemble(m2).download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.abs(a[2] - b[2]).max() <= 1e-12 * np.abs(a[2]).max()


def test_small_chunks_force_overflow_retry(ctx):
    # Hex20 vertex rows get 160 entries; 4-entry chunks overflow the 8-chunk
    # directory and the pipeline must retry with larger chunks.
    m = alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=4)
    out = ctx.assemble(m, chunk_slots=4)
    st = ctx.stats()
    gpu = out.download()
    assert st["retries"] >= 1 and st["chunk_slots"] >= 32
    compare_csr(gpu, oracle.assemble(m))
    # 2D rows with overflow chunks inside the directory (4 slots, 16 entries)
    m2 = alls.generate(alls.GEN_Q4, 30, 20, jitter=0.2, seed=4)
    compare_csr(ctx.assemble(m2, chunk_slots=4).download(), oracle.assemble(m2))


def fan_mesh(k):
    """k triangles around node 0: that row gets 3k list entries."""
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    coords = np.concatenate([[0.0, 0.0], np.column_stack([np.cos(ang), np.sin(ang)]).reshape(-1)])
    conn = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    return alls.MeshArrays(coords, conn, 2, 3, uniform_type=alls.T3)


@pytest.mark.parametrize("k", [40, 300, 700])
def test_big_row_fallback_for_unstructured_valence(ctx, k):
    # 120 entries: more than the 32-lane register path holds, so the row goes
    # to the LDS hash/bucket kernel; 900 entries: same kernel at its largest
    # LDS size; 2100 entries: beyond it, the LDS sort kernel.
    m = fan_mesh(k)
    out = ctx.assemble(m)
    assert ctx.stats()["big_rows"] >= 1
    compare_csr(out.download(), oracle.assemble(m))


def test_row_beyond_consolidation_limit_is_refused(ctx):
    # 1400 triangles on one node: 4200 entries in one node row, more than the
    # 4096 the consolidation holds in LDS -> a clear error before the write pass.
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble(fan_mesh(1400))
    assert ei.value.status == alls.E_UNSUPPORTED


def test_empty_rows_and_orphan_nodes(ctx):
    m = alls.generate(alls.GEN_Q4, 6, 5, jitter=0.1, seed=2)
    coords = np.concatenate([m.coords, [5.0, 5.0, 6.0, 6.0]])   # two unused nodes
    m2 = alls.MeshArrays(coords, m.conn, 2, 4)
    gpu = ctx.assemble(m2).download()
    assert gpu[0][-1] == gpu[0][-5]
    compare_csr(gpu, oracle.assemble(m2))


def test_zero_elements(ctx):
    m = alls.MeshArrays(np.zeros(8), np.zeros(0, np.int32), 2, 4)
    rp, ci, v = ctx.assemble(m).download()
    assert (rp == 0).all() and ci.size == 0


def test_invalid_node_id_is_reported(ctx):
    m = alls.generate(alls.GEN_Q4, 4, 4)
    conn = m.conn.copy()
    conn[7] = m.n_nodes + 3
    bad = alls.MeshArrays(m.coords, conn, 2, 4)
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble(bad)
    assert ei.value.status == alls.E_INVALID
    # the context stays usable
    compare_csr(ctx.assemble(m).download(), oracle.assemble(m))


def test_int64_row_ptr_option(ctx):
    m = alls.generate(alls.GEN_T3, 20, 20, jitter=0.2, seed=1, random_diagonal=1)
    out = ctx.assemble(m, row_ptr_i64=1)
    assert out.csr.row_ptr_is_i64 == 1
    compare_csr(out.download(), oracle.assemble(m))


@pytest.mark.parametrize("i64", [0, 1], ids=["rp32", "rp64"])
@pytest.mark.parametrize("kind,kw", [
    (alls.GEN_Q4, dict(nx=3, ny=3)),                          # 16 nodes: one whole run ending at n
    (alls.GEN_Q4, dict(nx=1, ny=14)),                         # 30 nodes: a whole run, then a partial one
    (alls.GEN_Q4, dict(nx=63, ny=63)),                        # 4096 nodes: exactly one scan tile
    (alls.GEN_Q4, dict(nx=64, ny=64, jitter=0.2, seed=5)),    # 4225 nodes: a second, partial tile
    (alls.GEN_HEX20, dict(nx=3, ny=2, nz=2, jitter=0.1, seed=6)),
], ids=["q4_16", "q4_30", "q4_4096", "q4_4225", "hex20_3x2x2"])
def test_row_ptr_scan_at_run_and_tile_edges(ctx, kind, kw, i64):
    """row_ptr is written 16 nodes per thread as 16-byte vectors; runs that
    end exactly at n, partial runs and partial tiles, both index widths."""
    print(f"case kind={kind} {kw} i64={i64}", flush=True)
    m, gpu = run_case(ctx, kind, kw, row_ptr_i64=i64)
    compare_csr(gpu, oracle.assemble(m))


def test_output_buffer_reuse(ctx):
    big = alls.generate(alls.GEN_Q4, 40, 40, jitter=0.2, seed=5)
    small = alls.generate(alls.GEN_T3, 10, 10, jitter=0.2, seed=6, random_diagonal=1)
    out = ctx.assemble(big)
    ptr = out.csr.val
    out = ctx.assemble(small, out=out)
    assert out.csr.val == ptr
    compare_csr(out.download(), oracle.assemble(small))


# ------------------------------------- full-size, size-independent checks ---
@pytest.mark.slow
def test_c2_q4_2000_full_size_invariants(ctx):
    n = 2000
    m = alls.generate(alls.GEN_Q4, n, n, jitter=0.2, seed=0xA11500C2, E_sigma=0.5)
    print("c2 generated", flush=True)
    rp, ci, v = ctx.assemble(m).download()
    assert rp[-1] == 4 * (3 * n + 1) ** 2 == 144048004
    # rows sorted and unique
    d = np.diff(ci.astype(np.int64))
    starts = rp[1:-1]
    d[starts[starts < d.size] - 1] = 1
    assert (d > 0).all()
    X = m.coords
    s = np.abs(v).max()
    for u in rigid_modes(X, 2):
        assert np.abs(csr_matvec(rp, ci, v, u)).max() < 1e-10 * s
    eps = np.array([[0.2, 0.15], [0.15, -0.1]])
    ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
    expect = (element_areas_2d(m) * m.E_per_elem).sum() * (ev @ plane_stress_D(1.0, m.nu) @ ev)
    u = uniform_strain_field(X, 2, eps)
    assert abs(u @ csr_matvec(rp, ci, v, u) - expect) < 1e-9 * abs(expect)


def test_matrix_market_file_round_trips_exactly(ctx, tmp_path):
    import scipy.io
    m = alls.generate(alls.GEN_MIXED, 9, 7, jitter=0.25, seed=17, random_diagonal=1, E_sigma=0.5)
    out = ctx.assemble(m)
    rp, ci, v = out.download()
    path = tmp_path / "K.mtx"
    out.write_mtx(path)
    out.free()
    K = scipy.io.mmread(str(path)).tocsr()
    K.sort_indices()
    assert K.shape == (rp.size - 1, rp.size - 1)
    assert np.array_equal(K.indptr, rp) and np.array_equal(K.indices, ci)
    assert np.array_equal(K.data, v)


@pytest.mark.parametrize("prec", [alls.PREC_F64, alls.PREC_F32_MFMA], ids=["f64_integration", "f32_mfma"])
def test_hex20_fp32_output_values(ctx, prec):
    # opts.val_f32: the reference's Hex20 path returns float values
    # (Hex:260, Hex:601-608).  Sums stay FP64 and only the final store narrows, so the
    # float output is exactly the FP64-output result of the same path rounded
    # to float; against the FP64 oracle it stays within 1e-6 Frobenius.
    m = alls.generate(alls.GEN_HEX20, 5, 4, 3, jitter=0.15, seed=14, E_sigma=0.5)
    out = ctx.assemble(m, precision=prec)
    rp64, ci64, v64 = out.download()
    out.free()
    out = ctx.assemble(m, precision=prec, val_f32=1)
    assert out.csr.val_is_f32 == 1
    rp, ci, v = out.download()
    out.free()
    assert np.array_equal(rp, rp64) and np.array_equal(ci, ci64)
    assert np.array_equal(v, v64.astype(np.float32).astype(np.float64))
    ref = oracle.assemble(m)[2]
    fro = float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
    print(f"hex20 float output, precision {prec}: Frobenius {fro:.3e}")
    assert fro <= 1e-6


def test_fp32_output_values_are_for_hex20_only(ctx):
    m = alls.generate(alls.GEN_Q4, 6, 5)
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble(m, val_f32=1)
    assert ei.value.status == alls.E_UNSUPPORTED
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble_batch([alls.generate(alls.GEN_HEX20, 2, 2, 2)], val_f32=1)
    assert ei.value.status == alls.E_UNSUPPORTED


T3_LAYOUT_CASES = [
    ("fused", lambda: alls.generate(alls.GEN_T3, 60, 50, jitter=0.3, seed=21, random_diagonal=1, permute_nodes=1,
                                    E_sigma=0.5), {}),
    ("overflow_chunks", lambda: alls.generate(alls.GEN_T3, 30, 20, jitter=0.3, seed=22, random_diagonal=1,
                                              permute_nodes=1), {"chunk_slots": 6}),
    ("fan_40", lambda: fan_mesh(40), {}),
    ("fan_300", lambda: fan_mesh(300), {}),
    ("fan_700", lambda: fan_mesh(700), {}),
    ("repeated_node", lambda: repeated_node_mesh(alls.GEN_T3, REPEATED[1][1])[1], {}),
]


@pytest.mark.parametrize("name,make,opts", T3_LAYOUT_CASES, ids=[c[0] for c in T3_LAYOUT_CASES])
def test_t3_line_groups_equal_keyed_lists(ctx, monkeypatch, name, make, opts):
    # Triangle meshes list each element row as one 128-byte line group and
    # merge rows with 32-bit keys; meshes with node ids past 27 bits key every
    # entry instead.  ALLS_T3_KEYED forces the keyed layout: both must give
    # the same bits (the serial element order) and match the oracle, through
    # the fused pass, overflow chunks, and the long-row kernels (fans).
    m = make()
    print(f"t3 layout case {name}", flush=True)
    line = ctx.assemble(m, **opts).download()
    monkeypatch.setenv("ALLS_T3_KEYED", "1")
    keyed = ctx.assemble(m, **opts).download()
    monkeypatch.delenv("ALLS_T3_KEYED")
    assert same_csr(line, keyed)
    compare_csr(line, oracle.assemble(m))


def test_t3_batch_line_groups_equal_keyed_lists(ctx, monkeypatch):
    meshes = [alls.generate(alls.GEN_T3, 12 + 3 * i, 9 + i, jitter=0.25, seed=40 + i, random_diagonal=1,
                            permute_nodes=1, E_sigma=0.5) for i in range(5)]
    out, ro, no = ctx.assemble_batch(meshes)
    line = out.download()
    out.free()
    monkeypatch.setenv("ALLS_T3_KEYED", "1")
    out, ro2, no2 = ctx.assemble_batch(meshes)
    keyed = out.download()
    out.free()
    monkeypatch.delenv("ALLS_T3_KEYED")
    assert np.array_equal(ro, ro2) and np.array_equal(no, no2)
    assert same_csr(line, keyed)


def _hex20_doubled(nx, ny, nz, seed, pick=None):
    # Elements listed again as new elements (all of them, or `pick`).
    m = alls.generate(alls.GEN_HEX20, nx, ny, nz, jitter=0.1, seed=seed, E_sigma=0.5)
    c = m.conn.reshape(m.n_elems, 20)
    idx = np.arange(m.n_elems) if pick is None else np.asarray(pick)
    conn = np.concatenate([c, c[idx]])
    E = None if m.E_per_elem is None else np.concatenate([m.E_per_elem, m.E_per_elem[idx]])
    return alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, E, m.nu, m.thickness)


def _hex20_wide_row():
    m = alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=12)
    conn = m.conn.reshape(m.n_elems, 20).copy()
    x = int(conn[0, 0])
    far = [e for e in range(m.n_elems) if x not in conn[e]][2::4][:6]
    conn[far, 7] = x
    return alls.MeshArrays(m.coords, conn, 3, 20, m.uniform_type, None, m.E, None, m.nu, m.thickness)


HEX20_LAYOUT_CASES = [
    ("brick", lambda: alls.generate(alls.GEN_HEX20, 5, 4, 3, jitter=0.15, seed=14, E_sigma=0.5), {}),
    ("f32", lambda: alls.generate(alls.GEN_HEX20, 4, 3, 3, jitter=0.15, seed=15, E_sigma=0.5),
     {"precision": alls.PREC_F32_MFMA}),
    ("f32_float_values", lambda: alls.generate(alls.GEN_HEX20, 4, 3, 3, jitter=0.15, seed=16),
     {"precision": alls.PREC_F32_MFMA, "val_f32": 1}),
    ("overflow_chunks", lambda: alls.generate(alls.GEN_HEX20, 3, 3, 3, jitter=0.1, seed=8), {"chunk_slots": 40}),
    ("rows_past_count_table", lambda: _hex20_doubled(2, 2, 2, 9), {}),
    ("four_register_slots", lambda: _hex20_doubled(2, 2, 2, 11, [0, 5]), {}),
    ("wide_row_deferred", _hex20_wide_row, {}),
    ("repeated_node", lambda: repeated_node_mesh(*REPEATED[3])[1], {}),
    ("single", lambda: alls.generate(alls.GEN_HEX20, 1, 1, 1), {}),
]


@pytest.mark.parametrize("name,make,opts", HEX20_LAYOUT_CASES, ids=[c[0] for c in HEX20_LAYOUT_CASES])
def test_hex20_segment_lists_equal_keyed_lists(ctx, monkeypatch, name, make, opts):
    # Hex20 lists carry one element id per 20-slot strip (columns from the
    # connectivity); ALLS_HEX20_KEYED forces a key per entry.  Both give the
    # same bits and match the oracle, through overflow chunks, the second
    # count launch, four register slots, rows the segment merge defers, and
    # an element that lists a node twice.
    m = make()
    print(f"hex20 layout case {name}", flush=True)
    seg = ctx.assemble(m, **opts).download()
    monkeypatch.setenv("ALLS_HEX20_KEYED", "1")
    keyed = ctx.assemble(m, **opts).download()
    monkeypatch.delenv("ALLS_HEX20_KEYED")
    assert same_csr(seg, keyed)
    ref = oracle.assemble(m)
    if opts.get("precision", alls.PREC_F64) == alls.PREC_F64:
        compare_csr(seg, ref)
    else:
        assert np.array_equal(seg[0], ref[0]) and np.array_equal(seg[1], ref[1])
        assert np.linalg.norm(seg[2] - ref[2]) / np.linalg.norm(ref[2]) <= 1e-6


def test_hex20_batch_segment_lists_equal_keyed_lists(ctx, monkeypatch):
    meshes = [alls.generate(alls.GEN_HEX20, 2 + i % 3, 2, 1 + i % 2, jitter=0.1, seed=50 + i, E_sigma=0.5)
              for i in range(5)]
    out, ro, no = ctx.assemble_batch(meshes)
    seg = out.download()
    out.free()
    monkeypatch.setenv("ALLS_HEX20_KEYED", "1")
    out, ro2, no2 = ctx.assemble_batch(meshes)
    keyed = out.download()
    out.free()
    monkeypatch.delenv("ALLS_HEX20_KEYED")
    assert np.array_equal(ro, ro2) and np.array_equal(no, no2)
    assert same_csr(seg, keyed)

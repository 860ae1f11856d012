"""GPU parity on randomised meshes: random kind, size, jitter, node
permutation, element subsets (which leave orphan nodes and ragged boundary
rows), element order and E fields.  Each mesh goes through the plain path
(every consolidation kernel: register rows, fused T3/mixed pass, LDS rows),
the cached-pattern path and one batch of all of them; all must agree with
the oracle (bit-exact pattern, values within 1e-12 * sum|contributions|) and
with each other bitwise."""
import numpy as np
import pytest

import alls
import oracle
from _checks import compare_csr

pytestmark = pytest.mark.gpu


def random_mesh(seed):
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(4))
    if kind == alls.GEN_HEX20:
        m = alls.generate(kind, int(rng.integers(1, 4)), int(rng.integers(1, 4)), int(rng.integers(1, 4)),
                          jitter=float(rng.uniform(0, 0.15)), seed=seed)
        npe = 20
    else:
        m = alls.generate(kind, int(rng.integers(2, 40)), int(rng.integers(2, 40)), jitter=float(rng.uniform(0, 0.3)),
                          seed=seed, random_diagonal=int(rng.integers(2)), permute_nodes=int(rng.integers(2)),
                          split_fraction=float(rng.uniform(0.1, 0.9)))
        npe = m.npe
    conn = m.conn.reshape(-1, npe)
    keep = rng.random(conn.shape[0]) < rng.uniform(0.5, 1.0)
    keep[0] = True
    order = rng.permutation(np.nonzero(keep)[0])
    et = m.elem_type[order] if m.elem_type is not None else None
    E = np.exp(0.5 * rng.standard_normal(order.size))
    return alls.MeshArrays(m.coords, conn[order].reshape(-1), m.dim, npe, m.uniform_type, et, 1.0, E, m.nu,
                           m.thickness)


@pytest.mark.parametrize("seed", range(24))
def test_random_mesh_all_paths(ctx, seed):
    m = random_mesh(1000 + seed)
    ref = oracle.assemble(m)
    out = ctx.assemble(m)
    got = out.download()
    out.free()
    compare_csr(got, ref)
    pat = alls.Pattern(ctx, m)
    vals = pat.values(m.E_per_elem)
    for a, b in zip(vals.download(), got):
        assert np.array_equal(a, b)
    vals.free()
    pat.free()


def test_random_meshes_in_batches(ctx):
    meshes = [random_mesh(2000 + s) for s in range(16)]
    for dim in (2, 3):
        group = [m for m in meshes if m.dim == dim]
        if not group:
            continue
        out, row_off, nnz_off = ctx.assemble_batch(group)
        parts = alls.split_batch(out.download(), row_off, nnz_off)
        out.free()
        for m, p in zip(group, parts):
            alone = ctx.assemble(m)
            for a, b in zip(p, alone.download()):
                assert np.array_equal(a, b)
            alone.free()


@pytest.mark.parametrize("seed", range(12))
def test_random_mesh_chunk_sizes_and_layouts(ctx, monkeypatch, seed):
    # Random chunk sizes (small ones give every row overflow chunks; Hex20
    # takes strips of 20 slots when the size is a multiple of 20, keyed
    # entries otherwise) and both list layouts of each kind (T3 / mixed line
    # groups or keys, Hex20 strips or keys): every variant gives the same
    # bits, and the oracle's matrix.
    m = random_mesh(3000 + seed)
    rng = np.random.default_rng(seed)
    if m.dim == 3:
        sizes = [0, int(rng.choice([20, 40, 60, 80])), int(rng.integers(21, 140))]
        env = "ALLS_HEX20_KEYED"
    else:
        sizes = [0, int(rng.integers(4, 13)), int(rng.integers(13, 64))]
        env = "ALLS_T3_KEYED"
    print(f"random mesh {3000 + seed}: dim {m.dim}, {m.n_elems} elements, chunk sizes {sizes}", flush=True)
    ref = oracle.assemble(m)
    first = None
    for keyed in (False, True):
        if keyed:
            monkeypatch.setenv(env, "1")
        for c in sizes:
            out = ctx.assemble(m, chunk_slots=c)
            got = out.download()
            out.free()
            if first is None:
                first = got
                compare_csr(got, ref)
            else:
                for a, b in zip(got, first):
                    assert np.array_equal(a, b), (keyed, c)
    monkeypatch.delenv(env)

"""Batched assembly (alls_assemble_batch) and the autotuner, through the C ABI.

A batch is assembled in one pipeline over the packed meshes; each mesh's
block of the block-diagonal result must be bit-identical to assembling that
mesh alone (same kernels, every column summed in element order), and within
the oracle tolerance of the CPU oracle.  The captured hipGraph (the whole
pipeline, one launch per call) is replayed
several times with new E fields and eager work in between; every replay is
checked.  Launch settings (autotuned or not) must not change any bit.
"""
import numpy as np
import pytest

import alls
import oracle
from _checks import compare_csr

pytestmark = pytest.mark.gpu


def mesh_set_2d():
    return [
        alls.generate(alls.GEN_Q4, 12, 9, jitter=0.2, seed=21, E_sigma=0.5),
        alls.generate(alls.GEN_T3, 7, 11, jitter=0.3, seed=22, random_diagonal=1, permute_nodes=1),
        alls.generate(alls.GEN_MIXED, 15, 4, jitter=0.25, seed=23, random_diagonal=1, E=2.5),
        alls.generate(alls.GEN_Q4, 1, 1),
        alls.generate(alls.GEN_Q4, 40, 33, jitter=0.2, seed=24, E_sigma=0.5),
    ]


def alone(ctx, m, **kw):
    out = ctx.assemble(m, **kw)
    r = out.download()
    out.free()
    return r


def check_batch(ctx, meshes, got, row_off, nnz_off, **kw):
    parts = alls.split_batch(got, row_off, nnz_off)
    assert len(parts) == len(meshes)
    for m, p in zip(meshes, parts):
        ref = alone(ctx, m, **kw)
        for a, b in zip(p, ref):
            assert np.array_equal(a, b)


def test_batch_2d_equals_each_mesh_alone_and_oracle(ctx):
    meshes = mesh_set_2d()
    out, row_off, nnz_off = ctx.assemble_batch(meshes)
    got = out.download()
    out.free()
    assert row_off[-1] == sum(m.n_dofs for m in meshes) and nnz_off[-1] == got[1].size
    check_batch(ctx, meshes, got, row_off, nnz_off)
    for m, p in zip(meshes, alls.split_batch(got, row_off, nnz_off)):
        compare_csr(p, oracle.assemble(m))


@pytest.mark.parametrize("prec", [alls.PREC_F64, alls.PREC_F32_MFMA])
def test_batch_hex20(ctx, prec):
    meshes = [alls.generate(alls.GEN_HEX20, 3, 2, 2, jitter=0.15, seed=31, E_sigma=0.5),
              alls.generate(alls.GEN_HEX20, 1, 1, 1),
              alls.generate(alls.GEN_HEX20, 4, 3, 5, jitter=0.1, seed=32)]
    out, row_off, nnz_off = ctx.assemble_batch(meshes, precision=prec)
    got = out.download()
    out.free()
    check_batch(ctx, meshes, got, row_off, nnz_off, precision=prec)
    if prec == alls.PREC_F64:
        for m, p in zip(meshes, alls.split_batch(got, row_off, nnz_off)):
            compare_csr(p, oracle.assemble(m))


def test_batch_device_meshes(ctx):
    meshes = mesh_set_2d()
    dms = [alls.DeviceMesh(m) for m in meshes]
    out, row_off, nnz_off = ctx.assemble_batch(meshes, c_meshes=[d.c_mesh() for d in dms])
    got = out.download()
    out.free()
    for d in dms:
        d.free()
    check_batch(ctx, meshes, got, row_off, nnz_off)


def test_batch_graph_replays_with_new_E(ctx):
    # Same shape every call: the pipeline is captured once as one graph and
    # replayed, one hipGraphLaunch per call.  Eager assemblies on the same
    # context run between the replays.
    rng = np.random.default_rng(5)
    meshes = mesh_set_2d()
    out = None
    for it in range(4):
        for m in meshes:
            m.E_per_elem = np.exp(0.5 * rng.standard_normal(m.n_elems))
        out, row_off, nnz_off = ctx.assemble_batch(meshes, out=out)
        assert ctx.stats()["graph_launches"] == 1
        got = out.download()
        check_batch(ctx, meshes, got, row_off, nnz_off)
        for m, p in zip(meshes[:2], alls.split_batch(got, row_off, nnz_off)[:2]):
            compare_csr(p, oracle.assemble(m))
    out.free()


def test_batch_rejects_mixed_dimensions(ctx):
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble_batch([alls.generate(alls.GEN_Q4, 2, 2), alls.generate(alls.GEN_HEX20, 1, 1, 1)])
    assert ei.value.status == alls.E_INVALID
    q = alls.generate(alls.GEN_Q4, 2, 2)
    q2 = alls.generate(alls.GEN_Q4, 2, 2, nu=0.25)
    with pytest.raises(alls.AllsError):
        ctx.assemble_batch([q, q2])


@pytest.mark.parametrize("kind,kw", [(alls.GEN_Q4, dict(nx=50, ny=40, jitter=0.2, seed=41, E_sigma=0.5)),
                                     (alls.GEN_T3, dict(nx=30, ny=30, jitter=0.3, seed=42, random_diagonal=1,
                                                        permute_nodes=1)),
                                     (alls.GEN_HEX20, dict(nx=3, ny=3, nz=2, jitter=0.15, seed=43))])
def test_launch_settings_do_not_change_bits(ctx, kind, kw):
    m = alls.generate(kind, **kw)
    ref = alone(ctx, m)
    for bt, g, c in [(64, 0, 8), (128, 8, 12), (256, 32, 48), (64, 4, 160)]:
        if kind == alls.GEN_HEX20 and c < 20:
            continue
        got = alone(ctx, m, block_threads=bt, grid_per_cu=g, chunk_slots=c)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    best = ctx.autotune(m)
    assert best["chunk_slots"] > 0 and best["block_threads"] in (64, 128, 256) and best["grid_per_cu"] >= 0
    got = alone(ctx, m, autotune=1)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_batch_row_longer_than_the_captured_write_kernels(ctx):
    # 300 triangles around one node: a 900-entry row, longer than the 256
    # entries the captured long-row kernels take.  The graph's write pass
    # flags it and the write runs again, uncaptured, with the real bound.
    ang = np.linspace(0, 2 * np.pi, 300, endpoint=False)
    coords = np.concatenate([[0.0, 0.0], np.column_stack([np.cos(ang), np.sin(ang)]).reshape(-1)])
    conn = np.array([[0, 1 + i, 1 + (i + 1) % 300] for i in range(300)], np.int32)
    fan = alls.MeshArrays(coords, conn, 2, 3, uniform_type=alls.T3)
    meshes = [alls.generate(alls.GEN_T3, 12, 9, jitter=0.2, seed=31, random_diagonal=1), fan,
              alls.generate(alls.GEN_T3, 7, 7, jitter=0.2, seed=32)]
    out, row_off, nnz_off = ctx.assemble_batch(meshes)
    got = out.download()
    out.free()
    check_batch(ctx, meshes, got, row_off, nnz_off)
    compare_csr(alls.split_batch(got, row_off, nnz_off)[1], oracle.assemble(fan))

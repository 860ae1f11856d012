"""GPU: alls_assemble_noalloc, the entry point that allocates no device
memory (list pool and consolidation scratch in the library's fixed device
array), against alls_assemble on the same inputs: row_ptr, col_ind and val
bitwise equal, on two calls in a row, the second with a smaller mesh (the
scratch keeps the first call's contents)."""
import numpy as np
import pytest

import alls

pytestmark = pytest.mark.gpu

CASES = [
    ("q4", alls.GEN_Q4, dict(nx=120, ny=90, jitter=0.2, seed=1, E_sigma=0.5), dict(nx=40, ny=30), {}),
    ("t3", alls.GEN_T3, dict(nx=100, ny=80, jitter=0.3, seed=2, random_diagonal=1, permute_nodes=1),
     dict(nx=30, ny=25), {}),
    ("mixed", alls.GEN_MIXED, dict(nx=100, ny=80, jitter=0.3, seed=3, random_diagonal=1, permute_nodes=1,
                                   E_sigma=0.5), dict(nx=30, ny=20), {}),
    ("hex20_f64", alls.GEN_HEX20, dict(nx=8, ny=7, nz=6, jitter=0.15, seed=4, E_sigma=0.5), dict(nx=3, ny=3, nz=2),
     dict(precision=alls.PREC_F64)),
    ("hex20_f32", alls.GEN_HEX20, dict(nx=8, ny=7, nz=6, jitter=0.15, seed=5), dict(nx=4, ny=2, nz=2),
     dict(precision=alls.PREC_F32_MFMA)),
]


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("name,kind,big_kw,small_kw,opts", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_assemble_twice(ctx, name, kind, big_kw, small_kw, opts):
    big = alls.generate(kind, **big_kw)
    small = alls.generate(kind, **dict(big_kw, **small_kw))
    print(f"noalloc case {name}: {big.n_elems} then {small.n_elems} elements", flush=True)
    ref_big = ctx.assemble(big, **opts).download()
    out = alls.CallerCsr(ctx, ref_big[0].size, ref_big[1].size)
    for m in (big, small):
        ref = ctx.assemble(m, **opts).download()
        dm = alls.DeviceMesh(m)
        ctx.assemble_noalloc(dm, out, **opts)
        assert out.csr.n_rows == ref[0].size - 1 and out.csr.nnz == ref[1].size
        assert same(out.download(), ref)
        dm.free()
    out.free()


def test_noalloc_int64_row_ptr(ctx):
    m = alls.generate(alls.GEN_T3, 30, 30, jitter=0.2, seed=6, random_diagonal=1)
    ref = ctx.assemble(m).download()
    out = alls.CallerCsr(ctx, ref[0].size, ref[1].size, row_ptr_i64=1)
    dm = alls.DeviceMesh(m)
    ctx.assemble_noalloc(dm, out)
    assert same(out.download(), ref)
    dm.free()
    out.free()


def test_noalloc_refusals(ctx):
    m = alls.generate(alls.GEN_Q4, 40, 30, jitter=0.2, seed=7)
    ref = ctx.assemble(m).download()
    dm = alls.DeviceMesh(m)
    # output too small: refused, with the sizes it needs
    out = alls.CallerCsr(ctx, ref[0].size, ref[1].size - 1)
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble_noalloc(dm, out)
    assert ei.value.status == alls.E_INVALID and out.csr.nnz == ref[1].size
    out.free()
    # host-resident mesh: refused (it would need staging memory)
    out = alls.CallerCsr(ctx, ref[0].size, ref[1].size)
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble_noalloc(dm, out, c_mesh=m.c_mesh())
    assert ei.value.status == alls.E_INVALID
    # lists beyond the fixed scratch: refused with a clear error
    huge = alls.generate(alls.GEN_Q4, 600, 600)
    dh = alls.DeviceMesh(huge)
    with pytest.raises(alls.AllsError) as ei:
        ctx.assemble_noalloc(dh, out)
    assert ei.value.status == alls.E_UNSUPPORTED and "scratch" in str(ei.value)
    dh.free()
    # and the path still works afterwards
    ctx.assemble_noalloc(dm, out)
    assert same(out.download(), ref)
    out.free()
    dm.free()


def test_noalloc_hex20_fp32_output_values(ctx):
    m = alls.generate(alls.GEN_HEX20, 6, 5, 4, jitter=0.15, seed=9)
    ref = ctx.assemble(m, precision=alls.PREC_F32_MFMA, val_f32=1).download()
    out = alls.CallerCsr(ctx, ref[0].size, ref[1].size, val_f32=1)
    dm = alls.DeviceMesh(m)
    ctx.assemble_noalloc(dm, out, precision=alls.PREC_F32_MFMA, val_f32=1)
    assert same(out.download(), ref)
    dm.free()
    out.free()

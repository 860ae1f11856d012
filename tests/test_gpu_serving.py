"""GPU: the continuous-batching server.  Every result in a concurrent mix
must equal the same request assembled alone (and the oracle), within the
parity tolerance of _checks.py; repeated mesh keys must hit the cached
pattern; admission must respect the HBM budget."""
import numpy as np
import pytest

import alls
import oracle
from _checks import compare_csr

pytestmark = pytest.mark.gpu


def mix_requests():
    rng = np.random.default_rng(4)
    base = [
        alls.generate(alls.GEN_Q4, 60, 40, jitter=0.2, seed=1),
        alls.generate(alls.GEN_T3, 40, 40, jitter=0.3, seed=2, random_diagonal=1, permute_nodes=1),
        alls.generate(alls.GEN_MIXED, 40, 30, jitter=0.3, seed=3, random_diagonal=1, permute_nodes=1),
        alls.generate(alls.GEN_HEX20, 4, 3, 3, jitter=0.15, seed=4),
    ]
    reqs = []
    for rep in range(3):
        for k, m in enumerate(base):
            reqs.append((k + 1, m.with_E(np.exp(0.5 * rng.standard_normal(m.n_elems)))))
    return reqs


def test_mix_equals_alone_and_oracle(ctx):
    reqs = mix_requests()
    with alls.Server(n_streams=4) as srv:
        tickets = [srv.submit(m, key=key) for key, m in reqs]
        results = []
        for t in tickets:
            csr, st = srv.wait(t)
            results.append((srv.download(csr), st))
            srv.release(csr)
    hits = [st["cache_hit"] for _, st in results]
    assert sum(hits) >= 4            # later repeats of each key use the cached pattern
    assert all(st["status"] == 0 for _, st in results)
    for (key, m), (got, st) in zip(reqs, results):
        ref = oracle.assemble(m)
        compare_csr(got, ref)
        alone = ctx.assemble(m).download()
        compare_csr(alone, ref)
        assert np.array_equal(got[0], alone[0]) and np.array_equal(got[1], alone[1])
        assert (np.abs(got[2] - alone[2]) <= 2e-12 * ref[3]).all()


def test_value_only_hits_are_bitwise_repeatable():
    m = alls.generate(alls.GEN_Q4, 50, 50, jitter=0.2, seed=9)
    E = np.exp(0.5 * np.random.default_rng(2).standard_normal(m.n_elems))
    with alls.Server(n_streams=2) as srv:
        srv.wait(srv.submit(m, key=7), keep_output=False)
        outs = []
        for _ in range(3):
            csr, st = srv.wait(srv.submit(m.with_E(E), key=7))
            assert st["cache_hit"] == 1
            outs.append(srv.download(csr)[2])
            srv.release(csr)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_small_budget_serializes_and_refuses_oversized():
    small = alls.generate(alls.GEN_Q4, 40, 40, seed=1)
    big = alls.generate(alls.GEN_HEX20, 12, 12, 12, seed=1)
    budget = 6 * 1024 * 1024   # one small request (about 3.5 MB) at a time
    with alls.Server(n_streams=4, hbm_budget_bytes=budget) as srv:
        with pytest.raises(alls.AllsError) as ei:
            srv.submit(big)
        assert ei.value.status == alls.E_BUSY
        ts = [srv.submit(small.with_E(np.full(small.n_elems, 1.0 + i))) for i in range(8)]
        stats = [srv.wait(t, keep_output=False)[1] for t in ts]
    assert all(s["status"] == 0 for s in stats)

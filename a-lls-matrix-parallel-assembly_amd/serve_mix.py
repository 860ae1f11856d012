"""Served-mix workload driver: Poisson arrivals of assembly requests.

Request i draws its element type uniformly from {Q4, T3, Hex20} and its size
log-uniformly in [min_elems, max_elems]; with probability `reuse` it repeats
an earlier mesh (same mesh key, so the server's cached pattern applies) with
a fresh lognormal(0, 0.5) E field.  All meshes and E fields are generated
before the clock starts.  Latency is submit -> completion as the server
measures it; throughput is total elements over the makespan.
"""
import math
import threading
import time

import numpy as np

import alls


def _mesh_for(kind, n_elems, seed):
    if kind == alls.HEX20:
        n = max(1, round(n_elems ** (1.0 / 3.0)))
        return alls.generate(alls.GEN_HEX20, n, n, n, jitter=0.15, seed=seed)
    if kind == alls.T3:
        n = max(1, round(math.sqrt(n_elems / 2.0)))
        return alls.generate(alls.GEN_T3, n, n, jitter=0.3, seed=seed, random_diagonal=1, permute_nodes=1)
    n = max(1, round(math.sqrt(n_elems)))
    return alls.generate(alls.GEN_Q4, n, n, jitter=0.2, seed=seed)


def build_requests(n_requests, seed=0xA11500C5, min_elems=1e3, max_elems=1e6, reuse=0.6, progress=None):
    rng = np.random.default_rng(seed)
    meshes, reqs = [], []
    for i in range(n_requests):
        if progress and i % 250 == 0:
            progress(f"generating request {i}/{n_requests}")
        if meshes and rng.random() < reuse:
            mid = int(rng.integers(len(meshes)))
        else:
            kind = int(rng.integers(3))
            size = math.exp(rng.uniform(math.log(min_elems), math.log(max_elems)))
            meshes.append(_mesh_for(kind, size, seed + 1 + len(meshes)))
            mid = len(meshes) - 1
        m = meshes[mid]
        E = np.exp(0.5 * rng.standard_normal(m.n_elems))
        reqs.append((mid, m.with_E(E)))
    return meshes, reqs


# Default per-element service times in ms (first assembly, cached-pattern
# repeat), used only to set the arrival rate; bench.py passes measured ones.
DEFAULT_COSTS = {"2d_full": 7e-7, "2d_hit": 5e-7, "hex20_full": 6e-5, "hex20_hit": 3e-5, "fixed": 0.2}


def expected_service_ms(m, first=True, costs=None):
    """Single-stream service-time estimate of one request."""
    c = costs or DEFAULT_COSTS
    key = ("hex20" if m.dim == 3 else "2d") + ("_full" if first else "_hit")
    return c["fixed"] + m.n_elems * c[key]


def offered_busy_s(reqs, costs=None):
    seen = set()
    total = 0.0
    for mid, m in reqs:
        total += expected_service_ms(m, mid not in seen, costs)
        seen.add(mid)
    return total / 1e3


def one_stream_ms(server, reqs):
    """Mean wall time (ms) per request when requests are served one at a
    time (submit, wait): the measured one-stream service time of the mix."""
    t0 = time.perf_counter()
    for mid, m in reqs:
        server.wait(server.submit(m, key=mid + 1), keep_output=False)
    return (time.perf_counter() - t0) * 1e3 / max(len(reqs), 1)


def run(server, reqs, rate_per_s, seed=1, progress=None):
    """Submits reqs at Poisson arrival times; returns per-request stats."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=len(reqs))
    arrivals = np.cumsum(gaps)
    tickets = [None] * len(reqs)
    stats = [None] * len(reqs)
    ready = threading.Semaphore(0)
    errors = []

    def submitter():
        t0 = time.perf_counter()
        for i, (mid, m) in enumerate(reqs):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            try:
                tickets[i] = server.submit(m, key=mid + 1)
            except Exception as e:   # refused request: record and go on
                errors.append((i, str(e)))
            ready.release()

    th = threading.Thread(target=submitter)
    t_start = time.perf_counter()
    th.start()
    for i in range(len(reqs)):
        ready.acquire()
        if tickets[i] is None:
            continue
        try:
            out, st = server.wait(tickets[i], keep_output=False)
            stats[i] = st
        except alls.AllsError as e:    # a failed request is recorded, the mix goes on
            errors.append((i, str(e)))
        if progress and (i + 1) % 200 == 0:
            progress(f"served {i + 1}/{len(reqs)} requests, {time.perf_counter() - t_start:.1f} s")
    th.join()
    makespan = time.perf_counter() - t_start
    return stats, makespan, errors


def summarize(reqs, stats, makespan, errors):
    lat = np.array([s["latency_ms"] for s in stats if s is not None])
    q = np.array([s["queue_ms"] for s in stats if s is not None])
    hits = sum(1 for s in stats if s is not None and s["cache_hit"] == 1)
    bsz = [s.get("batch_size", 1) for s in stats if s is not None]
    batched = sum(1 for b in bsz if b > 1)
    paths = {}
    for (_, m), st in zip(reqs, stats):
        if st is None:
            continue
        k = ("hex20" if m.dim == 3 else "2d") + "_" + ("hit", "plain", "build")[[1, 0, 2].index(st["cache_hit"])]
        d = paths.setdefault(k, {"n": 0, "elements": 0, "run_ms": 0.0})
        d["n"] += 1
        d["elements"] += m.n_elems
        d["run_ms"] += st["run_ms"]
    for d in paths.values():
        d["ns_per_element"] = round(d["run_ms"] * 1e6 / max(d["elements"], 1), 2)
        d["run_ms"] = round(d["run_ms"], 1)
    elems = sum(m.n_elems for (_, m), s in zip(reqs, stats) if s is not None)
    return {"requests": len(reqs), "completed": int(lat.size), "refused": len(errors),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "queue_p99_ms": round(float(np.percentile(q, 99)), 3),
            "cache_hits": hits, "batched_requests": batched,
            "mean_batch_size_of_batched": round(float(np.mean([b for b in bsz if b > 1])), 2) if batched else 0.0,
            "paths": paths, "elements": int(elems), "makespan_s": round(makespan, 3),
            "elements_per_s": round(elems / makespan, 1)}

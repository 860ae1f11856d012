"""Benchmark of the A-LLS assembly path on one MI355X.

One step = one full A-LLS assembly (integration + list insertion + device
consolidation into CSR) of the workload mesh, inputs already resident in
HBM.  The default workload is the Q4 plane-stress 2000x2000 mesh (4.0M
elements, 8.0M DOF, jitter +-0.2h, lognormal(0, 0.5) E per element).

Prints one JSON line.  Extra keys: roofline of the dominant kernel (hipEvent
times on the library's stream), the serial CPU baseline (the oracle on a
bounded sample of the same workload, one thread), and secondary
measurements (Hex20 FP64 vs FP32-MFMA, GPU coloring, serving) under "extra".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "a-lls-matrix-parallel-assembly_amd"))

METRIC = "elements/s & %HBM roofline vs GPU coloring; FP32-MFMA speedup; served p99 ms"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def q4_workload(n, seed=0xA11500C2):
    import alls
    return alls.generate(alls.GEN_Q4, n, n, jitter=0.2, seed=seed, E_sigma=0.5, nu=0.3, thickness=1.0)


def hex20_workload(n, seed=0xA11500C4):
    import alls
    return alls.generate(alls.GEN_HEX20, n, n, n, jitter=0.15, seed=seed, E=1.0, nu=0.3)


def c3_workload(name, n, seed=0xA11500C3):
    """Unstructured T3 ("t3": every cell split along a random diagonal) or
    mixed Q4/T3 ("mixed": 50% of cells split) on an n x n grid, jitter
    +-0.3h, node ids randomly permuted, E=1, nu=0.3."""
    import alls
    kind, kw = {"t3": (alls.GEN_T3, {}), "mixed": (alls.GEN_MIXED, {"split_fraction": 0.5})}[name]
    return alls.generate(kind, n, n, jitter=0.3, seed=seed, random_diagonal=1, permute_nodes=1, **kw)


def time_steps(ctx, dmesh, steps, warmup, stage_steps=5, **opts):
    """W untimed + K timed assemblies, then a few instrumented ones; returns
    (seconds of the K, their stats with stage times averaged over the
    instrumented runs, out).  The timed runs ask for no stage timing
    (opts.timing 0): the library then records no hipEvents between the
    stages, which would add a few microseconds of idle time at each.  The
    instrumented runs are the same pipeline with the events."""
    import alls
    cm = dmesh.c_mesh()
    out = None
    for _ in range(warmup):
        out = ctx.assemble(dmesh.mesh, out=out, c_mesh=cm, timing=0, **opts)
    alls.hip().sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.assemble(dmesh.mesh, out=out, c_mesh=cm, timing=0, **opts)
    alls.hip().sync()
    dt = time.perf_counter() - t0
    acc = {}
    n_inst = max(1, min(steps, stage_steps))
    for _ in range(n_inst):
        out = ctx.assemble(dmesh.mesh, out=out, c_mesh=cm, timing=1, **opts)
        st = ctx.stats()
        for k in ("ms_total", "ms_init", "ms_insert", "ms_count", "ms_scan", "ms_write"):
            acc[k] = acc.get(k, 0.0) + st[k]
    st = ctx.stats()
    for k in acc:
        st[k] = acc[k] / n_inst
    c = out.csr
    st["out_allocated_bytes"] = (c.cap_rows * (8 if c.row_ptr_is_i64 else 4)
                                 + c.cap_nnz * (4 + (4 if c.val_is_f32 else 8)))
    return dt, st, out


def path_bytes(mesh, nnz):
    """Compulsory bytes of one assembly: the mesh read once (coordinates,
    connectivity, per-element E) and the CSR written once (row_ptr 4 B per
    row + 1, col_ind 4 B and val 8 B per nonzero).  The list entries that
    A-LLS writes and reads back in between are intermediate traffic, not
    algorithmic bytes."""
    inputs = mesh.coords.nbytes + mesh.conn.nbytes + (mesh.E_per_elem.nbytes if mesh.E_per_elem is not None else 0)
    if mesh.elem_type is not None:
        inputs += mesh.elem_type.nbytes
    return inputs + 4 * (mesh.n_dofs + 1) + 12 * nnz


def intermediate_bytes(mesh, st, val_bytes, key_bytes):
    """List traffic per stage (bytes moved beyond the compulsory ones): the
    entries the insertion appends (values + keys or group ids) and the
    consolidation passes read back."""
    entries = st["list_entries"]
    return {"insert": int(entries * (val_bytes + key_bytes)),
            "count": int(entries * key_bytes),
            "write": int(entries * (val_bytes + key_bytes))}


def hbm_roofline(comp_bytes, stage_ms, stage_kernel):
    """Roofline of the dominant kernel: the compulsory bytes of one assembly
    (one launch of each kernel) over that kernel's average duration (hipEvents
    on the library's stream around the stage, which holds that one kernel)."""
    dom = max(stage_ms, key=stage_ms.get)
    achieved = comp_bytes / (stage_ms[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "stage": dom,
            "kernel": stage_kernel.get(dom, dom), "kernel_ms": round(stage_ms[dom], 4),
            "algorithmic_bytes_per_launch": comp_bytes}


def path_block(comp_bytes, ms):
    gbs = comp_bytes / (ms * 1e-3) / 1e9
    return {"compulsory_bytes": comp_bytes, "ms_per_step": round(ms, 4), "achieved_GBs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}


def reservations(mesh):
    """Reservation atomics of one insertion: one per (element, row node)
    before any merging of equal rows (Q4 4, T3 3, Hex20 20 per element)."""
    if mesh.elem_type is not None:
        nn = np.where(mesh.elem_type == 0, 4, 3)
        return int(nn.sum())
    return mesh.n_elems * {4: 4, 3: 3, 20: 20}[mesh.npe]


def merged_reservations(mesh):
    """Global reservation atomics the insertion issues: the 2D kernel merges
    the reservations of a wave (64 consecutive elements) by row node, so it
    issues one per distinct node of each wave; a Hex20 wave is one element,
    one per distinct node of the element."""
    c = mesh.conn.reshape(-1, mesh.npe)
    group = 1 if mesh.npe == 20 else 64
    pad = (-c.shape[0]) % group
    if pad:
        c = np.concatenate([c, np.full((pad, c.shape[1]), -1, np.int32)])
    c = np.sort(c.reshape(-1, group * mesh.npe), axis=1)
    distinct = (np.diff(c, axis=1) != 0) & (c[:, 1:] >= 0)
    return int(distinct.sum() + (c[:, 0] >= 0).sum())


def atomic_block(n_res, insert_ms, atomic_peak_gops, issued=None):
    """Reservation atomics per second of the insertion kernel against the
    device's measured rate of atomic adds to distinct addresses."""
    rate = n_res / (insert_ms * 1e-3) / 1e9
    out = {"reservations_per_launch": n_res, "reservations_per_s_G": round(rate, 2),
           "measured_atomic_Gops": round(atomic_peak_gops, 2),
           "frac_of_measured": round(rate / atomic_peak_gops, 4) if atomic_peak_gops else None}
    if issued is not None:
        r2 = issued / (insert_ms * 1e-3) / 1e9
        out.update({"atomics_issued_per_launch": issued, "atomics_issued_per_s_G": round(r2, 2),
                    "issued_frac_of_measured": round(r2 / atomic_peak_gops, 4) if atomic_peak_gops else None})
    return out


def footprint(mesh, st):
    """HBM held by one assembly: inputs, list headers + pool, output CSR (its
    nnz, and the buffers allocated for it: room for the list-entries bound,
    so that the pipeline needs one host synchronisation)."""
    inputs = path_bytes(mesh, 0) - 4 * (mesh.n_dofs + 1)
    rp = (mesh.n_dofs + 1) * (8 if st["nnz"] >= 2 ** 31 else 4)
    out = rp + 12 * st["nnz"]
    alloc = st.get("out_allocated_bytes", out)
    lists = st["scratch_bytes"]
    return {"inputs": inputs, "lists": lists, "output": out, "output_allocated": alloc,
            "total": inputs + lists + alloc}


# Kernel behind each pipeline stage of the Q4 workload (rocprofv3 names).
STAGE_KERNEL = {"insert": "k_insert_2d<0>", "count": "k_count_quad<4, double>",
                "write": "k_rows_small<16, 1, 4, 4, double, true, int, double>"}


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (same workload), already scaled by the
    copy-kernel calibration; None when the summary is absent."""
    path = os.path.join(REPO, "profiles", "alls_bench_hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["traffic_bytes"], t["source"]
    except (OSError, KeyError, ValueError):
        return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh):
    """Oracle (serial restatement of Acc:329-380 + sorted pattern) on the
    whole workload mesh, one thread pinned to one core."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        rp, ci = oracle.pattern(mesh)
        t1 = time.perf_counter()
        oracle.assemble(mesh, rp, ci, with_abs=False)
        t2 = time.perf_counter()
    finally:
        os.sched_setaffinity(0, old)
    return {"value": round(mesh.n_elems / (t2 - t0), 1), "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"the whole workload mesh ({mesh.n_elems} elements): sorted pattern {t1 - t0:.2f} s + "
                      f"serial element-order fill {t2 - t1:.2f} s, one thread pinned to core {core}",
            "seconds": round(t2 - t0, 2), "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def time_calls(fn, steps, warmup):
    import alls
    out = None
    for _ in range(warmup):
        out = fn(out)
    alls.hip().sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn(out)
    alls.hip().sync()
    return (time.perf_counter() - t0) / steps, out


def coloring_compare(ctx, mesh, dm, steps, alls_ms):
    """Cached-pattern value-only (per-row gather, no atomics) vs GPU colouring on the same mesh."""
    import alls
    alls.Pattern(ctx, mesh, c_mesh=dm.c_mesh()).free()   # warm-up
    t0 = time.perf_counter()
    pat = alls.Pattern(ctx, mesh, c_mesh=dm.c_mesh())
    t_build = time.perf_counter() - t0
    pat.out.free()
    Ep = dm.E_per_elem
    res = {"pattern_build_ms": round(t_build * 1e3, 3)}
    dt, out = time_calls(lambda o: pat.values(device_ptr=Ep, out=o), steps, 2)
    res["value_only_ms"] = round(dt * 1e3, 4)
    st = ctx.stats()
    res["value_only_device_ms"] = round(st["ms_total"], 4)
    out.free()
    dt, out = time_calls(lambda o: pat.values(E_per_elem=mesh.E_per_elem, out=o), steps, 1)
    res["value_only_host_E_ms"] = round(dt * 1e3, 4)
    out.free()
    for name, mk in (("structured_4color", lambda: pat.color_from_host(mesh.structured_color)),
                     ("gpu_jpl", lambda: pat.color_gpu())):
        t0 = time.perf_counter()
        col = mk()
        t_col = time.perf_counter() - t0
        dt, out = time_calls(lambda o: pat.colored(col, device_ptr=Ep, out=o), steps, 2)
        res[name] = {"n_colors": col.n_colors, "coloring_ms": round(t_col * 1e3, 2),
                     "colored_assembly_ms": round(dt * 1e3, 4),
                     "elements_per_s": round(mesh.n_elems / dt, 1),
                     "alls_full_vs_colored_speedup": round(dt * 1e3 / alls_ms, 3)}
        out.free()
        col.free()
    res["note"] = ("A-LLS full = integration + lists + device pattern + values; colored and value-only run on a "
                   "cached pattern and scatter map (numeric phase only), as the reference's colored timing excludes "
                   "its host pattern and colouring")
    pat.free()
    return res


# MFMA work of one Hex20 element in k_hex20: the 10 upper 16x16 tiles of the
# padded 64x64 product X^T (W X), 7 k-steps of 16x16x4 (2048 flops each).
HEX20_MFMA_FLOPS = 10 * 7 * 2 * 16 * 16 * 4
# The dense B^T D B contraction count per element (60 x 60 x 162 x 2).
HEX20_DENSE_FLOPS = 1166400


def hex20_compare(ctx, n, steps, peaks):
    """FP64 vs FP32-MFMA Hex20 assembly of an n^3 brick (one GPU)."""
    import alls
    m = hex20_workload(n)
    dm = alls.DeviceMesh(m)
    res = {}
    v = {}
    for name, prec in (("f64", alls.PREC_F64), ("f32_mfma", alls.PREC_F32_MFMA)):
        dt, st, out = time_steps(ctx, dm, steps, 1, precision=prec)
        ms = dt / steps * 1e3
        res[name] = {"ms_per_step": round(ms, 3), "elements_per_s": round(m.n_elems * steps / dt, 1),
                     "ms_insert": round(st["ms_insert"], 3), "ms_count": round(st["ms_count"], 3),
                     "ms_write": round(st["ms_write"], 3), "nnz": st["nnz"]}
        times = {"insert": st["ms_insert"], "count": st["ms_count"], "write": st["ms_write"]}
        t = "double" if name == "f64" else "float"
        kernels = {"insert": f"k_hex20<{t}, 0>", "count": "k_rows_count_long<4, true>",
                   "write": f"k_rows_merge_seg<{t}, int, 3, 20, double, true>"}
        pb = path_bytes(m, st["nnz"])
        res[name]["roofline"] = hbm_roofline(pb, times, kernels)
        res[name]["roofline"]["traffic"] = measured_traffic(res[name]["roofline"]["kernel"])[0]
        res[name]["path"] = path_block(pb, ms)
        peak = FP64_PEAK_TFLOPS if name == "f64" else FP32_PEAK_TFLOPS
        meas = peaks["mfma_f64_TFLOPs"] if name == "f64" else peaks["mfma_f32_TFLOPs"]
        real = m.n_elems * HEX20_MFMA_FLOPS / (st["ms_insert"] * 1e-3) / 1e12
        res[name]["mfma"] = {
            "kernel": kernels["insert"], "kernel_ms": round(st["ms_insert"], 4),
            "flops_per_element": HEX20_MFMA_FLOPS, "achieved_TFLOPs": round(real, 2), "peak_TFLOPs": peak,
            "utilisation": round(real / peak, 4), "measured_peak_TFLOPs": round(meas, 2),
            "frac_of_measured": round(real / meas, 4),
            "dense_equivalent_TFLOPs": round(m.n_elems * HEX20_DENSE_FLOPS / (st["ms_insert"] * 1e-3) / 1e12, 1),
            "note": "MFMA flops the kernel issues (10 tiles x 7 k-steps of 16x16x4 per element) over its time; "
                    "dense_equivalent counts the B^T D B contraction (1,166,400 flops per element), which the "
                    "M-form does not perform"}
        res[name]["atomics"] = atomic_block(reservations(m), st["ms_insert"], peaks["atomic_f64_Gops"],
                                            merged_reservations(m))
        res[name]["hbm_footprint"] = footprint(m, st)
        v[name] = out.download()[2]
        dt_h, out = time_calls(lambda o: ctx.assemble(m, precision=prec, out=o), 2, 1)
        res[name]["host_input_ms"] = round(dt_h * 1e3, 3)
        out.free()
    # The reference's Hex20 path returns float values: FP32-MFMA integration
    # with FP32 output values (sums still FP64, only the stores narrow).
    dt, st, out = time_steps(ctx, dm, steps, 1, precision=alls.PREC_F32_MFMA, val_f32=1)
    ms = dt / steps * 1e3
    r32 = {"ms_per_step": round(ms, 3), "elements_per_s": round(m.n_elems * steps / dt, 1),
           "ms_insert": round(st["ms_insert"], 3), "ms_count": round(st["ms_count"], 3),
           "ms_write": round(st["ms_write"], 3), "speedup_vs_f64": round(res["f64"]["ms_per_step"] / ms, 3)}
    if v["f64"] is not None:
        vf = out.download()[2]
        r32["frobenius_vs_f64"] = float(np.linalg.norm(vf - v["f64"]) / np.linalg.norm(v["f64"]))
        del vf
    out.free()
    res["f32_mfma_f32_values"] = r32
    # Cached pattern: value-only re-assembly (staged, no atomics) and the
    # coloured path with the 8-colour cell-parity colouring.
    for name, prec in (("f64", alls.PREC_F64), ("f32_mfma", alls.PREC_F32_MFMA)):
        t0 = time.perf_counter()
        pat = alls.Pattern(ctx, m, precision=prec, c_mesh=dm.c_mesh())
        res[name]["pattern_build_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        pat.out.free()
        dt, out = time_calls(lambda o: pat.values(out=o), steps, 1)
        res[name]["value_only_ms"] = round(dt * 1e3, 3)
        Eh = np.full(m.n_elems, 1.0)
        dt, out = time_calls(lambda o: pat.values(E_per_elem=Eh, out=o), 2, 1)
        res[name]["value_only_host_E_ms"] = round(dt * 1e3, 3)
        if name == "f64":
            col = pat.color_from_host(m.structured_color)
            dtc, out2 = time_calls(lambda o: pat.colored(col, out=o), steps, 1)
            res["colored_8color_f64_ms"] = round(dtc * 1e3, 3)
            res["alls_full_vs_colored_speedup_f64"] = round(dtc * 1e3 / res["f64"]["ms_per_step"], 3)
            out2.free()
            col.free()
        out.free()
        pat.free()
    dm.free()
    res["speedup_f32_vs_f64"] = round(res["f64"]["ms_per_step"] / res["f32_mfma"]["ms_per_step"], 3)
    if v["f64"] is not None:
        res["frobenius_f32_vs_f64"] = float(np.linalg.norm(v["f32_mfma"] - v["f64"]) / np.linalg.norm(v["f64"]))
    res["workload"] = f"Hex20 {n}^3 brick, jitter +-0.15h, E=1, nu=0.3, 3x3x3 Gauss"
    return res


def c3_compare(ctx, n, steps, peaks):
    """Unstructured T3 and mixed Q4/T3 n x n (jitter 0.3h, random diagonals,
    permuted node ids): A-LLS full path, value-only and GPU-coloured paths."""
    import alls
    res = {}
    for name in ("t3", "mixed"):
        m = c3_workload(name, n)
        dm = alls.DeviceMesh(m)
        dt, st, out = time_steps(ctx, dm, steps, 1)
        out.free()
        ms = dt / steps * 1e3
        r = {"elements": m.n_elems, "nnz": st["nnz"], "ms_per_step": round(ms, 3),
             "elements_per_s": round(m.n_elems * steps / dt, 1), "ms_insert": round(st["ms_insert"], 3),
             "ms_count": round(st["ms_count"], 3), "ms_write": round(st["ms_write"], 3), "big_rows": st["big_rows"]}
        times = {"insert": st["ms_insert"], "count": st["ms_count"], "write": st["ms_write"]}
        # T3 lists are line groups (3 value blocks + a 32-byte key block per
        # 128-byte line, NG 3); mixed lists key every entry (8 bytes).
        # T3 takes the fused pass; mixed meshes count and write separately.
        if name == "t3":
            mode, ng = 1, 3
            kernels = {"insert": "k_insert_2d<1>", "write": "k_rows_fused<32, 3, int>"}
        else:
            mode, ng = 2, 0
            kernels = {"insert": "k_insert_2d<2>", "count": "k_rows_small<32, 2, 0, 4, double, false, int, double>",
                       "write": "k_rows_small<32, 1, 0, 4, double, true, int, double>"}
        pb = path_bytes(m, st["nnz"])
        r["roofline"] = hbm_roofline(pb, times, kernels)
        r["roofline"]["traffic"] = measured_traffic(r["roofline"]["kernel"])[0]
        r["path"] = path_block(pb, ms)
        r["intermediate_bytes"] = intermediate_bytes(m, st, 4 * 8, 32 / 3 if ng == 3 else 8)
        r["atomics"] = atomic_block(reservations(m), st["ms_insert"], peaks["atomic_f64_Gops"],
                                    merged_reservations(m))
        r["hbm_footprint"] = footprint(m, st)
        pat = alls.Pattern(ctx, m, c_mesh=dm.c_mesh())
        pat.out.free()
        dtv, o2 = time_calls(lambda o: pat.values(out=o), steps, 1)
        r["value_only_ms"] = round(dtv * 1e3, 3)
        o2.free()
        col = pat.color_gpu()
        dtc, o3 = time_calls(lambda o: pat.colored(col, out=o), steps, 1)
        r["gpu_jpl_colors"] = col.n_colors
        r["colored_ms"] = round(dtc * 1e3, 3)
        r["alls_full_vs_colored_speedup"] = round(dtc / (dt / steps), 3)
        o3.free()
        col.free()
        pat.free()
        dm.free()
        res[name] = r
    res["workload"] = (f"{n}x{n} grid, jitter +-0.3h, random diagonal per cell, node ids randomly permuted; "
                       "T3: every cell split; mixed: 50% of cells split; E=1, nu=0.3, FP64")
    return res


def batch_compare(ctx, n_meshes=256, seed=7):
    """Many small meshes: one alls_assemble_batch call (packed pipeline,
    persistent insertion, replayed hipGraphs) vs one alls_assemble per mesh."""
    import alls
    rng = np.random.default_rng(seed)
    meshes = []
    for i in range(n_meshes):
        kind = (alls.GEN_Q4, alls.GEN_T3, alls.GEN_MIXED)[i % 3]
        size = float(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
        nx = max(2, int(np.sqrt(size / (2.0 if kind == alls.GEN_T3 else 1.0))))
        meshes.append(alls.generate(kind, nx, nx, jitter=0.2, seed=seed + i, random_diagonal=1, E_sigma=0.5))
    n_el = sum(m.n_elems for m in meshes)
    dms = [alls.DeviceMesh(m) for m in meshes]
    cms = [d.c_mesh() for d in dms]
    out = None
    for _ in range(2):   # first call captures the graphs
        out, _, _ = ctx.assemble_batch(meshes, out=out, c_meshes=cms)
    alls.hip().sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        out, _, _ = ctx.assemble_batch(meshes, out=out, c_meshes=cms)
    t_batch = (time.perf_counter() - t0) / reps
    out.free()
    outs = [None] * n_meshes
    for k in range(n_meshes):
        outs[k] = ctx.assemble(meshes[k], out=outs[k], c_mesh=cms[k])
    t0 = time.perf_counter()
    for _ in range(reps):
        for k in range(n_meshes):
            outs[k] = ctx.assemble(meshes[k], out=outs[k], c_mesh=cms[k])
    t_seq = (time.perf_counter() - t0) / reps
    for o in outs:
        o.free()
    for d in dms:
        d.free()
    return {"meshes": n_meshes, "elements": n_el, "batch_ms": round(t_batch * 1e3, 3),
            "one_by_one_ms": round(t_seq * 1e3, 3), "batch_elements_per_s": round(n_el / t_batch, 1),
            "speedup": round(t_seq / t_batch, 2),
            "workload": f"{n_meshes} 2D meshes (Q4, T3, mixed in turn), 1e3-3e4 elements log-uniform, device-resident"}


def autotune_report(ctx, dmesh, steps, default_ms):
    """alls_autotune on the main workload and the tuned setting's time."""
    import alls
    t0 = time.perf_counter()
    best = ctx.autotune(dmesh.mesh, c_mesh=dmesh.c_mesh())
    t_tune = time.perf_counter() - t0
    dt, st, out = time_steps(ctx, dmesh, steps, 1, **best)
    out.free()
    return {"setting": best, "tuning_s": round(t_tune, 2), "tuned_ms_per_step": round(dt / steps * 1e3, 4),
            "default_ms_per_step": round(default_ms, 4)}


def serve_costs_from(res, extra, n_q4):
    """Per-element service times (ms) measured earlier in this run, with the
    inputs in host memory as the served requests have them."""
    import serve_mix
    c = dict(serve_mix.DEFAULT_COSTS)
    c["2d_full"] = res.get("host_input", {}).get("ms_per_step", res["ms_per_step"]) / n_q4
    cq = extra.get("coloring_q4", {})
    if "value_only_ms" in cq:
        c["2d_hit"] = cq.get("value_only_host_E_ms", cq["value_only_ms"]) / n_q4
    hx = extra.get("hex20", {})
    if "f64" in hx:
        f = hx["f64"]
        n_hex = f["elements_per_s"] * f["ms_per_step"] / 1e3
        c["hex20_full"] = f.get("host_input_ms", f["ms_per_step"]) / n_hex
        c["hex20_hit"] = f.get("value_only_host_E_ms", f.get("value_only_ms", f["ms_per_step"])) / n_hex
    return c


def served_mix(n_requests, costs, streams=3):
    """Poisson mix of Q4/T3/Hex20 requests (1e3-1e6 elements, 60% mesh reuse).
    The arrival rate puts the offered load at 75% of one stream's measured
    service rate: a separate sample of 200 requests from the same
    distribution is served one at a time first, by a server with one worker
    (so the rate does not depend on the served configuration).  The per-element cost model
    (`costs`, from this run's large-mesh measurements) is reported beside it;
    it underestimates small and concurrent requests."""
    import alls
    import serve_mix
    t0 = time.perf_counter()
    meshes, reqs = serve_mix.build_requests(n_requests, progress=log)
    _, cal = serve_mix.build_requests(200, seed=0xA11500C6)
    t_gen = time.perf_counter() - t0
    with alls.Server(n_streams=1) as srv:
        per_req_ms = serve_mix.one_stream_ms(srv, cal)
    rate = 0.75 * 1e3 / per_req_ms
    busy_s = serve_mix.offered_busy_s(reqs, costs)
    with alls.Server(n_streams=streams) as srv:
        stats, makespan, errors = serve_mix.run(srv, reqs, rate, progress=log)
        counters = srv.counters()
    res = serve_mix.summarize(reqs, stats, makespan, errors)
    res["server_counters"] = {
        "builds": {"2d": counters["builds"][0], "hex20": counters["builds"][1]},
        "hits": {"2d": counters["hits"][0], "hex20": counters["hits"][1]},
        "evictions": {"2d": counters["evictions"][0], "hex20": counters["evictions"][1]},
        "evicted_unused": {"2d": counters["evicted_unused"][0], "hex20": counters["evicted_unused"][1]},
        "batched": counters["batched"], "passed_by_express_lane": counters["passed"]}
    res.update({"arrival_rate_per_s": round(rate, 1), "unique_meshes": len(meshes),
                "one_stream_ms_per_request": round(per_req_ms, 3),
                "model_busy_s": round(busy_s, 2), "cost_model_ms_per_element": costs,
                "generation_s": round(t_gen, 1), "streams": streams,
                "workload": "Poisson arrivals; type uniform over {Q4, T3, Hex20}; size log-uniform 1e3-1e6 "
                            "elements; 60% reuse a seen mesh with a new lognormal(0,0.5) E (cached pattern)"})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2000, help="Q4 grid size (workload n x n)")
    ap.add_argument("--cpu-n", type=int, default=0,
                    help="CPU baseline on a Q4 n x n sample instead of the whole workload (0: the workload)")
    ap.add_argument("--hex-n", type=int, default=80)
    ap.add_argument("--c3-n", type=int, default=1500, help="T3 / mixed grid size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--extras", default="coloring,autotune,batch,c3,hex20,serve",
                    help="comma list of secondary measurements to run (or 'none')")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-streams", type=int, default=3, help="server workers (the last is the express lane)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="end with one device-to-device copy of BYTES (a copy kernel of known size, the "
                         "scale for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    ex = set() if a.no_extra or a.extras == "none" else set(a.extras.split(","))
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)

    import alls
    log(f"generating Q4 {a.n}x{a.n}")
    mesh = q4_workload(a.n)
    ctx = alls.Context(0)
    dm = alls.DeviceMesh(mesh)
    log("timing A-LLS steps")
    dt, st, out = time_steps(ctx, dm, a.steps, a.warmup)
    ms = dt / a.steps * 1e3
    value = mesh.n_elems * a.steps / dt
    times = {"insert": st["ms_insert"], "count": st["ms_count"], "write": st["ms_write"]}
    pb = path_bytes(mesh, st["nnz"])
    roof = hbm_roofline(pb, times, STAGE_KERNEL)
    roof["traffic"], traffic_src = measured_traffic(roof["kernel"])
    peaks = ctx.device_peaks()
    log(f"{value:.3e} elements/s, {ms:.3f} ms/step, stages {times}")
    out.free()
    dt_h, out = time_calls(lambda o: ctx.assemble(mesh, out=o), max(3, a.steps // 2), 1)
    out.free()
    res = {
        "metric": METRIC, "value": round(value, 1), "unit": "elements/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"A-LLS assembly, Q4 plane stress {a.n}x{a.n} elements ({mesh.n_elems} elements, "
                               f"{mesh.n_dofs} DOF, nnz {st['nnz']}), jitter +-0.2h, E~lognormal(0,0.5), "
                               f"nu=0.3, t=1, 2x2 Gauss, FP64, inputs resident in HBM",
                   "n_elems": mesh.n_elems, "n_dofs": mesh.n_dofs, "nnz": st["nnz"]},
        "roofline": dict(roof,
                         measured_copy_GBs=round(peaks["copy_GBs"], 1),
                         frac_of_measured_copy=round(roof["achieved"] / peaks["copy_GBs"], 4),
                         measured_note="alls_device_peaks on the same device: a 1 GiB 16-byte streaming copy, "
                                       "bytes read + written per second",
                         traffic_source=traffic_src,
                         note="achieved = compulsory bytes of one assembly (mesh read once + CSR written once, "
                              f"{pb / mesh.n_elems:.1f} B/element) / the dominant kernel's average duration; "
                              "traffic = that kernel's measured HBM bytes per launch, which include the list "
                              "entries it reads back (intermediate_bytes)"),
        "stages_ms": {k: round(v, 4) for k, v in times.items()},
        "device_peaks": {k: round(v, 2) for k, v in peaks.items()},
        "path": path_block(pb, ms),
        "intermediate_bytes": intermediate_bytes(mesh, st, 4 * 8, 1),
        "atomics": atomic_block(reservations(mesh), st["ms_insert"], peaks["atomic_f64_Gops"],
                                merged_reservations(mesh)),
        "hbm_footprint": footprint(mesh, st),
        "host_input": {"ms_per_step": round(dt_h * 1e3, 4), "elements_per_s": round(mesh.n_elems / dt_h, 1),
                       "note": "same workload with the mesh in host memory: pinned staging and the PCIe "
                               "copy inside the timed region (not the metric's value)"},
    }
    extra = {}
    if "coloring" in ex:
        log("coloring / value-only comparison")
        try:
            extra["coloring_q4"] = coloring_compare(ctx, mesh, dm, a.steps, ms)
        except Exception as e:
            extra["coloring_q4"] = {"error": str(e)}
    if "autotune" in ex:
        try:
            log("autotune on the main workload")
            extra["autotune_q4"] = autotune_report(ctx, dm, a.steps, ms)
        except Exception as e:
            extra["autotune_q4"] = {"error": str(e)}
    dm.free()
    if "batch" in ex:
        try:
            log("batched small meshes")
            extra["batch"] = batch_compare(ctx)
        except Exception as e:
            extra["batch"] = {"error": str(e)}
    if "c3" in ex:
        try:
            log(f"T3 / mixed {a.c3_n}x{a.c3_n}")
            extra["c3"] = c3_compare(ctx, a.c3_n, 5, peaks)
        except Exception as e:
            extra["c3"] = {"error": str(e)}
    if not a.no_cpu:
        cm = mesh if a.cpu_n in (0, a.n) else q4_workload(a.cpu_n)
        log(f"cpu baseline on {cm.n_elems} elements")
        res["cpu_baseline"] = cpu_baseline(cm)
        res["cpu_baseline"]["gpu_speedup"] = round(value / res["cpu_baseline"]["value"], 1)
    if "hex20" in ex:
        try:
            log(f"Hex20 {a.hex_n}^3 FP64 vs FP32-MFMA")
            extra["hex20"] = hex20_compare(ctx, a.hex_n, 3, peaks)
        except Exception as e:  # report, do not hide the main line
            extra["hex20"] = {"error": str(e)}
    ctx.close()   # its buffers go to the allocation cache, which the server hands back to the
                  # device before it takes its HBM budget
    if "serve" in ex and a.serve_requests > 0:
        try:
            log(f"served mix: {a.serve_requests} requests")
            costs = dict(serve_costs_from(res, extra, mesh.n_elems), fixed=0.2)
            extra["served_mix"] = served_mix(a.serve_requests, costs, a.serve_streams)
        except Exception as e:
            extra["served_mix"] = {"error": str(e)}
    if extra:
        res["extra"] = extra
    if a.calibrate_copy > 0:
        h = alls.hip()
        src, dst = h.malloc(a.calibrate_copy), h.malloc(a.calibrate_copy)
        h.dtod(dst, src, a.calibrate_copy)
        h.sync()
        h.free(src)
        h.free(dst)
        res["calibration_copy_bytes"] = a.calibrate_copy
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

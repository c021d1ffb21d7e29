"""Benchmark: MPIR solves/s at 1e-10 relative residual on the 512 x 512 x 16
board, Q1 apply GDoF/s, per-solve latency, the roofline of the solve's
dominant kernel (the fused Jacobi-PCG pass) and of the Q1 apply, and the
reference algorithm timed on the host's cores.

One step is one complete MPIR solve (FP64 outer residual, update and
convergence check; FP32 Jacobi-PCG inner solves) of the seeded board
(``--size`` x ``--size`` x ``--slabs``, copper/FR-4) with its right-hand side
already resident in HBM.  Beside the headline the default line carries, on
short schedules, the 200-board uneven batch, the 64-excitation DC board and
the served stream at 50 requests/s, each against the single-solve path on
the same work.  Prints one JSON line; ``--workload`` selects one of the
workloads as a line of its own.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT / "pcb-analysis_amd" / "src", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _time_launches(torch, stream, launch, launches: int) -> float:
    """Average seconds per launch, HIP events recorded on the launch stream."""

    for _ in range(3):
        launch()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    stream.synchronize()
    start.record(stream)
    for _ in range(launches):
        launch()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / 1e3 / launches


def copy_bandwidth(torch, stream, nbytes: int = 1 << 30) -> float:
    """Stream-copy GB/s of the box (read + write bytes over time)."""

    from pcbhip import _capi

    n = nbytes // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    lib = _capi.load()
    per = _time_launches(torch, stream, lambda: _capi.check(
        lib.pcbhip_copy_dev(src.data_ptr(), dst.data_ptr(), n, stream.cuda_stream)), 20)
    return 2 * nbytes / per / 1e9


def stream_ceilings(torch, stream, n: int, shapes=((1, 1), (3, 1), (6, 2), (6, 4), (9, 4), (13, 4))) -> list:
    """GB/s of ``reads`` + ``writes`` arrays of ``n`` floats streamed at once
    (``pcbhip_streams_dev``), HIP events on ``stream``: the ceiling of a
    kernel that moves that many streams.  The fused PCG pass reads three
    record arrays (state, node operator, element coefficients) and writes
    one: the (3, 1) row."""

    import ctypes

    from pcbhip import _capi

    lib = _capi.load()
    n = -(-n // 4) * 4
    pool = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(max(r + w for r, w in shapes))]
    for t in pool:
        t.fill_(1.0)
    rows = []
    for reads, writes in shapes:
        ins = (ctypes.c_void_p * reads)(*[pool[k].data_ptr() for k in range(reads)])
        outs = (ctypes.c_void_p * max(writes, 1))(*[pool[reads + k].data_ptr() for k in range(writes)] or [None])
        per = _time_launches(torch, stream, lambda: _capi.check(
            lib.pcbhip_streams_dev(ins, reads, outs, writes, n, stream.cuda_stream)), 30)
        gbs = (reads + writes) * n * 4 / per / 1e9
        rows.append(dict(reads=reads, writes=writes, us=per * 1e6, gbs=gbs, frac=gbs / HBM_PEAK_GBS))
        log(f"streams {reads} reads + {writes} writes of {n:,} floats: {per * 1e6:.1f} us, {gbs:.0f} GB/s")
    return rows


def profiled_traffic(element_shape, rows: int, zb: int, board: str = "standard"):
    """HBM bytes per launch of the FP32 apply at this tile shape, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of that board (standard
    or general; copy-calibrated), or None when that shape was not profiled."""

    s, r, c = element_shape
    stem = "board" if board == "standard" else board
    path = ROOT / "profiles" / f"q1hex_apply_f32_{stem}_{c}x{r}x{s}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(f"traffic_bytes rows={rows} zb={zb}:"):
            return int(line.split(":", 1)[1].split("#")[0].strip())
    return None


def profiled_pass_traffic(element_shape, rows: int, target: int, board: str = "standard"):
    """HBM bytes per launch of the fused PCG pass at this tile shape, from the
    committed rocprofv3 summary of that board (standard or general), or None."""

    s, r, c = element_shape
    stem = "board" if board == "standard" else board
    path = ROOT / "profiles" / f"q1hex_pcg_pass_{stem}_{c}x{r}x{s}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(f"traffic_bytes pass rows={rows} target={target}:"):
            return int(line.split(":", 1)[1].split("#")[0].strip())
    return None


def skipped_bytes(operator) -> int:
    """Bytes of the algorithmic count that the stencils legitimately skip on
    this operator: the Robin word (4 B) of every node of a plain node plane
    (all nodes free, Robin exactly +0) and the coefficient rows (12 B per
    element) of every uniform element layer, among the first 64 of each (the
    device's own criteria)."""

    s, r, c = operator.mesh.element_grid_shape
    per_plane, per_layer = (r + 1) * (c + 1), r * c
    robin = np.asarray(operator._robin_total_high).reshape(s + 1, per_plane)
    free = np.asarray(operator.free_nodes).reshape(s + 1, per_plane)
    plain = [bool(free[z].all() and (robin[z] == 0.0).all() and not np.signbit(robin[z]).any())
             for z in range(min(s + 1, 64))]
    coef = np.asarray(operator._coefficients_high).reshape(3, s, per_layer)
    uniform = [bool(all((coef[d, ez] == coef[d, ez, 0]).all() and not np.signbit(coef[d, ez]).any() for d in range(3)))
               for ez in range(min(s, 64))]
    return 4 * per_plane * sum(plain) + 12 * per_layer * sum(uniform)


def pass_roofline(operator, batch, rows: int, target: int, launches: int = 100, board: str = "standard") -> dict:
    """The dominant kernel of the solve, the fused Jacobi-PCG pass, timed with
    HIP events on the batch's stream (``pcbhip_batch_time_pass``)."""

    n = operator.size
    s, r, c = operator.mesh.element_grid_shape
    e = s * r * c
    per_us = batch.time_pass(launches)
    # Reads r, q, p, 1/diag, Robin and c (4 B/node each) and the coefficients
    # (12 B/element); writes p, q, r and c (4 B/node each).
    algorithmic = 40 * n + 12 * e
    achieved = algorithmic / (per_us * 1e-6) / 1e9
    traffic = profiled_pass_traffic(operator.mesh.element_grid_shape, rows, target, board)
    needed = algorithmic - skipped_bytes(operator)
    return dict(
        kernel=f"k_inner_step<float,{rows}> (fused FP32 Jacobi-PCG pass), target {target} workgroups",
        per_launch_us=per_us,
        roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                      traffic=traffic, algorithmic_bytes=algorithmic, hbm_gbs=_hbm_rate(traffic, per_us * 1e-6),
                      board_needed_bytes=needed, board_needed_gbs=needed / (per_us * 1e-6) / 1e9,
                      board_needed_frac=needed / (per_us * 1e-6) / 1e9 / HBM_PEAK_GBS),
    )


def _hbm_rate(traffic, seconds: float):
    """GB/s of the measured HBM bytes per launch (the profile's traffic),
    beside ``achieved``, which counts the algorithmic bytes: on boards with
    plain planes and uniform layers the kernels skip many of those."""

    return None if traffic is None else traffic / seconds / 1e9


def apply_roofline(operator, torch, stream, launches: int = 100, board: str = "standard") -> dict:
    """Tile-shape sweep of the FP32 Q1 apply (the autotuner's candidates),
    timed with HIP events on the stream it runs on; the fastest shape is kept."""

    from pcbhip import autotune

    n = operator.size
    s, r, c = operator.mesh.element_grid_shape
    e = s * r * c
    dev = operator.device_operator
    x = torch.randn(n, dtype=torch.float32, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    y = torch.empty_like(x)
    table = []
    for rows, zb in autotune.apply_candidates(operator.mesh.element_grid_shape):
        dev.set_shape(rows, zb)
        per = _time_launches(torch, stream, lambda: dev.apply_f32_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream),
                             launches)
        table.append((per, rows, zb))
        log(f"apply rows={rows} zb={zb}: {per * 1e6:.1f} us")
    per, rows, zb = min(table)
    dev.set_shape(rows, zb)
    traffic = profiled_traffic(operator.mesh.element_grid_shape, rows, zb, board)
    algorithmic = 13 * n + 12 * e  # x, y, robin (4 B/node each), free mask (1 B/node), coefficients (12 B/element)
    achieved = algorithmic / per / 1e9
    needed = algorithmic - skipped_bytes(operator)
    return dict(
        kernel=f"k_apply<float,{rows},0,0> (FP32 Q1 apply), {zb} planes per z block",
        rows=rows,
        zb=zb,
        per_launch_us=per * 1e6,
        gdof_s=n / per / 1e9,
        sweep=[dict(rows=r_, zb=z_, us=p_ * 1e6) for p_, r_, z_ in table],
        roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                      traffic=traffic, algorithmic_bytes=algorithmic, hbm_gbs=_hbm_rate(traffic, per),
                      board_needed_bytes=needed, board_needed_gbs=needed / per / 1e9,
                      board_needed_frac=needed / per / 1e9 / HBM_PEAK_GBS),
    )


def _oracle_error(op, threads: int, rhs, gpu_solution) -> tuple[dict, np.ndarray]:
    """The FP64 oracle's solve of the GPU's right-hand side (FP64 inner PCG,
    relative residual 1e-12) and the GPU solution's error against it,
    relative to the largest value of the oracle's rise."""

    t0 = time.perf_counter()
    x64, res64 = op.solve_mpir(rhs, rtol=1e-12, inner_f64=True, max_outer=60, max_inner=100000, threads=threads)
    seconds = time.perf_counter() - t0
    log(f"cpu fp64 oracle: {res64['outer_iterations']} outer / {res64['inner_iterations']} inner, {seconds:.1f} s")
    scale = float(np.max(np.abs(x64)))
    return dict(fp64_oracle_rel_error_gpu=float(np.max(np.abs(gpu_solution - x64))) / scale,
                fp64_oracle_converged=bool(res64["converged"]), fp64_oracle_solve_s=seconds), x64


def _iteration_sample(op, threads: int) -> tuple[float, float]:
    """Seconds per FP32 Jacobi-PCG iteration and per FP64 apply of the port on
    ``threads`` threads: the difference of a 2- and a 6-iteration PCG (it
    drops the per-call set-up) and two FP64 applies."""

    rhs = np.random.default_rng(0).standard_normal(op.size)
    op.pcg_f32(rhs, 1e-30, 1, threads=threads)
    t0 = time.perf_counter()
    _, short, _, _ = op.pcg_f32(rhs, 1e-30, 2, threads=threads)
    t1 = time.perf_counter()
    _, long, _, _ = op.pcg_f32(rhs, 1e-30, 6, threads=threads)
    per_iteration = ((time.perf_counter() - t1) - (t1 - t0)) / max(long - short, 1)
    t0 = time.perf_counter()
    for _ in range(2):
        op.apply_f64(rhs, threads=threads)
    return per_iteration, (time.perf_counter() - t0) / 2


# The reference's published two-level solve of its 4-slab fixture at 50 x 50
# elements (13,005 nodes), C++ on a Xeon Platinum 8581C, construction
# excluded (docs/THERMAL_MPIR_FEM.md:472-474).
FIXTURE_PUBLISHED = {"elements": 50, "nodes": 13005, "outer": 7, "inner": 584, "ms_1_thread": 216.5,
                     "ms_16_threads": 26.6}


def _fixture_calibration(threads: int) -> dict:
    """The port's two-level solve of the reference's 4-slab fixture, timed as
    the reference times it (median of 5 after a warm-up), on one thread and
    on ``threads``, beside the published figures."""

    from oracle import native as N
    from oracle import problem as P
    from pcbhip import workloads

    p = P.from_problem(workloads.four_slab_fixture(FIXTURE_PUBLISHED["elements"]))
    ops = P.operator_arrays(p)
    op = N.OracleOperator(ops)
    rhs = P.build_rhs(ops, P.default_reference(ops, p), N.OracleOperator(ops, stiffness=True).apply_f64)
    t0 = time.perf_counter()
    coarse = op.two_level()
    construction_ms = (time.perf_counter() - t0) * 1e3
    out = {"published": FIXTURE_PUBLISHED, "construction_ms": construction_ms}
    for team in sorted({1, threads}):
        result, ms = _timed_ms(lambda: op.solve_mpir(rhs, rtol=1e-10, max_outer=16, max_inner=200, threads=team,
                                                     coarse=coarse)[1], repeats=5, warmups=1)
        out[f"{team}_threads"] = {"solve_ms": ms, "outer": result["outer_iterations"],
                                  "inner": result["inner_iterations"], "converged": result["converged"]}
    out["ratio_1_thread_vs_published"] = out["1_threads"]["solve_ms"] / FIXTURE_PUBLISHED["ms_1_thread"]
    log(f"cpu fixture calibration: {out}")
    return out


def cpu_baseline(problem, rhs, gpu_solution, *, oracle_error: bool = True, calibrate: bool = True) -> dict:
    """The reference algorithm on the host's cores: the oracle's C
    restatement of hex_q1.cpp (FP32 Jacobi-PCG inner, FP64 outer, the GPU
    line's limits), built for this host with -march=native, solving the
    GPU's right-hand side whole on every thread the process may use.  Beside
    it: a one-thread figure (per-iteration sample on one thread, scaled by
    the whole solve's own counts), the port's two-level solve of the
    reference's 4-slab fixture against the published times, and the FP64
    oracle's solve, against which the GPU's and the port's answers are
    stated."""

    from oracle import native as N
    from oracle import problem as P

    build = N.use_native_build()
    op = N.OracleOperator(P.operator_arrays(P.from_problem(problem)))
    threads = N.threads_default()
    t0 = time.perf_counter()
    x_port, res = op.solve_mpir(rhs, rtol=1e-10, max_outer=60, max_inner=100000, threads=threads)
    port_s = time.perf_counter() - t0
    log(f"cpu reference algorithm: {port_s:.1f} s on {threads} threads, {res['outer_iterations']} outer / "
        f"{res['inner_iterations']} inner")
    per_iteration, per_f64 = _iteration_sample(op, 1)
    one_s = res["inner_iterations"] * per_iteration + res["high_operator_applications"] * per_f64
    out = dict(
        value=1.0 / port_s, unit="solves/s", cores=threads, kind="port",
        sample=f"the whole solve of the GPU's right-hand side on {threads} host threads (oracle C restatement of "
               f"hex_q1.cpp, {build}): {res['outer_iterations']} outer / {res['inner_iterations']} inner, "
               f"{port_s:.1f} s",
        build=build, converged=bool(res["converged"]), outer_iterations=res["outer_iterations"],
        inner_iterations=res["inner_iterations"],
        one_thread={"solves_per_s": 1.0 / one_s, "est_solve_s": one_s, "ms_per_iteration": per_iteration * 1e3,
                    "ms_per_fp64_apply": per_f64 * 1e3,
                    "sample": "4 FP32 Jacobi-PCG iterations and 2 FP64 applies on one thread, scaled by the "
                              "whole solve's counts"},
    )
    if calibrate:
        out["fixture_calibration"] = _fixture_calibration(threads)
    if oracle_error:
        errors, x64 = _oracle_error(op, threads, rhs, gpu_solution)
        scale = float(np.max(np.abs(x64)))
        out.update(errors, fp64_oracle_rel_error_cpu_reference=float(np.max(np.abs(x_port - x64))) / scale)
    return out


def _oracle_operator(front):
    """The oracle port's operator for a front-end operator: a thermal one from
    its problem, a hexahedral DC one from its mesh's coefficients and fixed
    nodes (no Robin term)."""

    from oracle import native as N
    from oracle import problem as P

    if hasattr(front, "problem"):
        return N.OracleOperator(P.operator_arrays(P.from_problem(front.problem)))
    coef = np.stack(front.mesh.element_coefficients())
    unit = P.unit_matrices()
    free = np.asarray(front.free_nodes).reshape(-1)
    robin = np.zeros(front.size)
    return N.OracleOperator(dict(element_shape=front.mesh.element_grid_shape, coef=coef, unit=unit,
                                 robin_total=robin, free=free,
                                 diag=P.diagonal(coef, unit, robin, free, front.mesh.node_shape)))


def cpu_node_iteration_cost(pairs, threads: int) -> dict:
    """Whole reference-algorithm solves (the oracle's C port: FP32 Jacobi-PCG
    inner, FP64 outer, the GPU's limits) of ``(front operator, rhs)`` pairs
    on ``threads`` host threads: their time, and the seconds per node-
    iteration (nodes x inner iterations) that scale it to a whole workload."""

    seconds, work, iterations = 0.0, 0, 0
    for front, rhs in pairs:
        op = _oracle_operator(front)
        t0 = time.perf_counter()
        _, res = op.solve_mpir(np.asarray(rhs, dtype=np.float64), rtol=1e-10, max_outer=60, max_inner=100000,
                               threads=threads)
        seconds += time.perf_counter() - t0
        work += op.size * max(res["inner_iterations"], 1)
        iterations += res["inner_iterations"]
    return dict(solves=len(pairs), seconds=seconds, inner_iterations=iterations, node_iterations=work,
                seconds_per_node_iteration=seconds / work, threads=threads)


def batch_leg(args, torch, *, count: int, warmup: int, steps: int) -> dict:
    """``count`` uneven boards solved together in one batch (every launch
    covers every board still iterating; tile shape from the autotuner) and
    the same boards one at a time through the single-solve path."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from pcbhip import _capi, autotune, workloads
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    log(f"batch: building {count} uneven boards")
    boards = workloads.uneven_boards(count)
    t0 = time.perf_counter()
    ops = [MatrixFreeThermalOperator(b, preconditioner=args.preconditioner) for b in boards]
    log(f"batch: operators ({args.preconditioner}) built in {time.perf_counter() - t0:.1f} s")
    rhss = [o.build_rhs(o.default_reference_temperature()) for o in ops]
    nodes = sum(o.size for o in ops)
    config = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000).to_c()
    batch = _capi.Batch([o.device_operator for o in ops])
    batch.set_schedule(compact_percent=args.compact_percent)
    slots = [(b, None, config) for b in rhss]
    for i, b in enumerate(rhss):
        batch.set(i, b, None, config)
    key = autotune.board_set_key("batch", [o.mesh.element_grid_shape for o in ops])
    shape = autotune.tune_batch(batch, key, slots=slots)
    log(f"batch: tile shape rows={shape['rows']} target={shape['target']} (cached {shape['cached']})")
    for _ in range(warmup):
        batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    batched_s = (time.perf_counter() - t0) / steps
    results = [batch.result(i, solution=False)[1] for i in range(count)]
    converged = sum(r["converged"] for r in results)
    work = sum(r["inner_iterations"] * _pass_bytes(o.device_operator) for r, o in zip(results, ops))
    log(f"batch: {batched_s * 1e3:.1f} ms for {count} solves ({converged} converged), "
        f"{work / batched_s / 1e9:.0f} GB/s of pass bytes")
    singles = [_capi.Batch([o.device_operator]) for o in ops]
    for sb, b in zip(singles, rhss):
        sb.set_shape(args.rows, args.target)
        sb.set(0, b, None, config)
        sb.run()  # warm-up: graph capture of every single-solve batch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sb in singles:
        sb.run()
    single_s = time.perf_counter() - t0
    log(f"batch: one at a time {single_s * 1e3:.1f} ms")
    # The batched answers are the single answers bit for bit.
    same = all(batch.result(i)[0].tobytes() == singles[i].result(0)[0].tobytes() for i in range(0, count, 17))
    for sb in singles:
        sb.close()
    batch.close()
    cpu = None
    if not args.no_cpu:
        from oracle.native import threads_default

        # Six boards spread over the batch, solved whole on the host; the
        # whole batch's host time is their cost per node-iteration times the
        # batch's node-iterations.
        sample = cpu_node_iteration_cost([(ops[i], rhss[i]) for i in range(0, count, max(1, count // 6))][:6],
                                         threads_default())
        total_work = sum(o.size * max(r["inner_iterations"], 1) for o, r in zip(ops, results))
        cpu_s = total_work * sample["seconds_per_node_iteration"]
        cpu = dict(sample, solves_per_s=count / cpu_s, est_batch_s=cpu_s,
                   sample_note="whole solves of 6 of the boards on the host, scaled by node-iterations")
        log(f"batch: cpu sample {sample['seconds']:.1f} s, est. {cpu_s:.1f} s for the batch")
    return {
        "workload": f"{count} boards, sides log-uniform 32..384, depth 4..16, {nodes:,} nodes in total",
        "cpu": cpu,
        "solves_per_s": count / batched_s,
        "ms_per_batch": batched_s * 1e3,
        "pass_gbs": work / batched_s / 1e9,
        "single_solve_path_pass_gbs": work / single_s / 1e9,
        "single_solve_path_solves_per_s": count / single_s,
        "speedup_vs_single_solve_path": single_s / batched_s,
        "converged": converged,
        "bitwise_equal_alone_and_in_batch": bool(same),
        "tile_shape": {k: shape[k] for k in ("rows", "target", "us", "cached")},
        "preconditioner": args.preconditioner,
        "total_inner_iterations": int(sum(r["inner_iterations"] for r in results)),
        "max_inner_iterations": int(max(r["inner_iterations"] for r in results)),
    }


def plain_fp64_leg(args, torch, *, steps: int = 10) -> dict:
    """The 64 x 64 x 8 board as plain FP64 Jacobi PCG to 1e-10: on the device
    (FP64 inner solve, one refinement step), and the same algorithm on the
    host (the oracle's C port, FP64 inner PCG), whole solves on every thread
    and on one."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from pcbhip import _capi, workloads
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    problem = workloads.board_64x64x8()
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    rhs = operator.build_rhs(operator.default_reference_temperature())
    plain = MPIRConfig(relative_tolerance=1e-10, inner_relative_tolerance=1e-10, max_outer_iterations=4,
                       max_inner_iterations=20000)
    batch = _capi.Batch([operator.device_operator], inner_fp64=True)
    batch.set(0, rhs, None, plain.to_c())
    batch.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    gpu_s = (time.perf_counter() - t0) / steps
    x, result = batch.result(0)
    batch.close()
    out = {"workload": f"board 64x64x8 elements ({operator.size:,} nodes), plain FP64 Jacobi PCG to 1e-10",
           "solves_per_s": 1.0 / gpu_s, "ms_per_solve": gpu_s * 1e3, "converged": result["converged"],
           "inner_iterations": result["inner_iterations"], "relative_residual": result["relative_residual"]}
    if not args.no_cpu:
        from oracle.native import threads_default

        op = _oracle_operator(operator)
        for team in (threads_default(), 1):
            (x_cpu, res), ms = _timed_ms(lambda: op.solve_mpir(rhs, rtol=1e-10, inner_tol=1e-10, max_outer=4,
                                                               max_inner=20000, inner_f64=True, threads=team),
                                         repeats=3, warmups=1)
            out[f"cpu_{team}_threads"] = {"solves_per_s": 1e3 / ms, "ms_per_solve": ms,
                                          "inner_iterations": res["inner_iterations"]}
        out["rel_difference_gpu_vs_cpu"] = float(np.max(np.abs(x - x_cpu)) / np.max(np.abs(x_cpu)))
    log(f"plain fp64 64x64x8: {out}")
    return out


def run_batch200(args, torch) -> dict:
    """The batch of uneven boards as its own line (``--workload batch200``)."""

    leg = batch_leg(args, torch, count=args.count, warmup=args.warmup, steps=args.steps)
    return {
        "metric": f"MPIR solves/s at 1e-10 rel. residual ({args.count} uneven boards, one batch)",
        "value": leg["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded uneven-board generator, seed 4)",
        "config": dict(leg, parallelism="single GPU"),
    }


def multirhs_leg(args, torch, *, warmup: int, steps: int, variants=("slots", "block_mfma", "block_vector")) -> dict:
    """64 terminal-pair excitations of the 256 x 256 x 12 copper/dielectric
    DC board on one device-resident operator: as plain slots of one batch
    (the default), with the block apply (groups of right-hand sides sharing
    the operator's reads; matrix-core or vector element action) and one at a
    time through the single-solve path."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from electrical.matrix_free_mpir_fem.hex_dc import MatrixFreeHexDCOperator
    from pcbhip import _capi, workloads

    board = workloads.dc_board(size=args.dc_size, slabs=12, pairs=64)
    operator = MatrixFreeHexDCOperator(board.mesh, reference_node=board.reference_node)
    rhss = [operator.build_rhs(t) for t in board.excitations]
    config = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000).to_c()
    plans = {"slots": (0, 0), "block_mfma": (args.block, args.engine), "block_vector": (args.block, 0)}

    def timed_batch(group: int, engine: int):
        batch = _capi.Batch([operator.device_operator] * len(rhss))
        batch.set_shape(args.rows if group == 0 else 4, args.target)
        batch.set_block(group, engine)
        for i, b in enumerate(rhss):
            batch.set(i, b, None, config)
        for _ in range(warmup):
            batch.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run()
        seconds = (time.perf_counter() - t0) / steps
        results = [batch.result(i, solution=False)[1] for i in range(len(rhss))]
        return batch, seconds, results

    runs = {}
    for name in variants:
        batch, seconds, results = timed_batch(*plans[name])
        runs[name] = (batch, seconds, results)
        log(f"64 rhs, {name}: {seconds * 1e3:.1f} ms, {sum(r['converged'] for r in results)} converged, "
            f"max inner {max(r['inner_iterations'] for r in results)}")
    single = _capi.Batch([operator.device_operator])
    single.set_shape(args.rows, args.target)
    single.set(0, rhss[0], None, config)
    single.run()
    t0 = time.perf_counter()
    for b in rhss:
        single.set(0, b, None, None)
        single.run()
    single_s = time.perf_counter() - t0
    single.close()
    log(f"64 rhs one at a time: {single_s * 1e3:.1f} ms")
    plain_batch, plain_s, plain_results = runs["slots"]
    cpu = None
    if not args.no_cpu:
        from oracle.native import threads_default

        sample = cpu_node_iteration_cost([(operator, rhss[0])], threads_default())
        cpu = dict(sample, solves_per_s=1.0 / sample["seconds"],
                   sample_note="one of the 64 excitations solved whole on the host")
        log(f"64 rhs: cpu {sample['seconds']:.1f} s per solve")
    out = {
        "cpu": cpu,
        "workload": f"{board.mesh.size:,} nodes, 64 terminal pairs in one copper cluster",
        "path": "plain FP32 slots of one batch sharing the device operator",
        "solves_per_s": len(rhss) / plain_s,
        "ms_per_batch": plain_s * 1e3,
        "single_solve_path_solves_per_s": len(rhss) / single_s,
        "speedup_vs_single_solve_path": single_s / plain_s,
        "converged": int(sum(r["converged"] for r in plain_results)),
        "total_inner_iterations": int(sum(r["inner_iterations"] for r in plain_results)),
        "max_inner_iterations": int(max(r["inner_iterations"] for r in plain_results)),
    }
    # The block apply's answers against the plain FP32 slots (both refined
    # to 1e-10 in FP64): relative difference in potential on the terminals'
    # copper cluster, where the 1e-10 residual determines it (floating
    # islands couple only through the 1e-12 S/m dielectric).
    top = board.mesh.node_shape[0] - 1
    on = np.zeros(board.mesh.node_shape[1:], dtype=bool)
    for dy in (0, 1):
        for dx in (0, 1):
            on[dy:dy + board.cluster.shape[0], dx:dx + board.cluster.shape[1]] |= board.cluster
    for name in variants:
        if name == "slots":
            continue
        batch, seconds, results = runs[name]
        diff = 0.0
        for i in range(0, len(rhss), 9):
            a = batch.result(i)[0].reshape(board.mesh.node_shape)[top][on]
            b = plain_batch.result(i)[0].reshape(board.mesh.node_shape)[top][on]
            diff = max(diff, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
        out[name] = {"solves_per_s": len(rhss) / seconds, "vs_plain_slots": plain_s / seconds,
                     "group": args.block, "engine": plans[name][1],
                     "converged": int(sum(r["converged"] for r in results)),
                     "max_rel_difference_vs_fp32_slots_on_cluster": diff}
    for batch, _, _ in runs.values():
        batch.close()
    return out


def run_multirhs64(args, torch) -> dict:
    """The 64-excitation DC board as its own line (``--workload multirhs64``)."""

    leg = multirhs_leg(args, torch, warmup=args.warmup, steps=args.steps)
    return {
        "metric": "MPIR solves/s at 1e-10 rel. residual (64 RHS, 256x256x12 DC board)",
        "value": leg["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded DC board generator, seed 3)",
        "config": dict(leg, parallelism="single GPU"),
    }


def run_apply_engines(args, torch) -> dict:
    """The standalone FP32 Q1 apply of the benchmark board (``--board
    general``: the board with no stencil shortcuts) with each element action:
    FP32 vector FMAs, and the matrix cores with a two- or three-term bf16
    split (best tile shape of each, HIP-event timing)."""

    from pcbhip import autotune
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    problem = _board(args)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    dev = operator.device_operator
    n = operator.size
    s, r, c = operator.mesh.element_grid_shape
    e = s * r * c
    stream = torch.cuda.Stream()
    x = torch.randn(n, dtype=torch.float32, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    y = torch.empty_like(x)
    engines = {}
    for engine in (0, 2, 3):
        dev.set_engine(engine)
        best = None
        for rows, zb in autotune.apply_candidates(operator.mesh.element_grid_shape):
            dev.set_shape(rows, zb)
            per = _time_launches(torch, stream,
                                 lambda: dev.apply_f32_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream), 100)
            if best is None or per < best[0]:
                best = (per, rows, zb)
        per, rows, zb = best
        # Matrix-core work of one apply, from the launch geometry: every wave
        # strip of every z block sweeps its element layers, and each layer
        # takes (rows + 1) element rows x 2 products (lanes 0..31, 32..63) x
        # (terms - 1) 32x32x16 bf16 products.
        entry = {"us": per * 1e6, "rows": rows, "zb": zb, "gdof_s": n / per / 1e9,
                 "algorithmic_gbs": (13 * n + 12 * e) / per / 1e9}
        if engine:
            nsx = -(-(c + 1) // 62)
            waves = nsx * -(-(r + 1) // rows)
            nsz = -(-(s + 1) // zb)
            layers = sum(min(s - 1, (k + 1) * zb - 1) - max(0, k * zb - 1) + 1 for k in range(nsz))
            products = waves * layers * (rows + 1) * 2 * (engine - 1)
            flops = products * 32 * 32 * 16 * 2
            entry["mfma_tflops"] = flops / per / 1e12
            entry["mfma_frac_of_bf16_dense_peak"] = flops / per / 2.5e15
        engines[str(engine)] = entry
        log(f"engine {engine}: {entry}")
    dev.set_engine(0)
    best = engines["0"]
    # The box's copy moving the apply's algorithmic bytes (also the byte
    # counters' calibration kernel in a profiled run).
    copy = copy_bandwidth(torch, stream, (13 * n + 12 * e) // 2 // 16 * 16)
    board = "general board" if args.board == "general" else "board"
    return {
        "metric": f"Q1 apply GDoF/s (FP32, {c}x{r}x{s} {board}) per element action",
        "value": best["gdof_s"],
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": 100,
        "warmup": 3,
        "ms_per_step": best["us"] / 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded board generator, seed 2), x ~ N(0, 1)",
        "config": {"workload": f"{board} {c}x{r}x{s} elements ({n:,} nodes)", "engines": engines,
                   "copy_gbs_at_apply_bytes": copy, "parallelism": "single GPU"},
    }


def _board(args):
    """The benchmark board, or with ``--board general`` the board on which no
    stencil shortcut applies (every plane has Robin terms, no layer is uniform)."""

    from pcbhip import workloads

    if args.board == "general":
        return workloads.general_board(size=args.size, slabs=args.slabs)
    return workloads.board_512x512x16(size=args.size, slabs=args.slabs)


def run_pass(args, torch) -> dict:
    """The fused Jacobi-PCG pass and the FP32 apply of one board over tile
    shapes: one MPIR solve, then each (rows, target) shape of the pass timed
    with HIP events on the batch's stream (``pcbhip_batch_time_pass``), the
    apply's candidate shapes on a torch stream, and the box's copy at the
    pass's algorithmic bytes."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from pcbhip import _capi
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    problem = _board(args)
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    n = operator.size
    s, r, c = operator.mesh.element_grid_shape
    e = s * r * c
    rhs = operator.build_rhs(operator.default_reference_temperature())
    batch = _capi.Batch([operator.device_operator])
    batch.set_shape(args.rows, args.target)
    batch.set(0, rhs, None, MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60,
                                       max_inner_iterations=100000).to_c())
    t0 = time.perf_counter()
    batch.run()
    solve_s = time.perf_counter() - t0
    x, result = batch.result(0)
    digest = hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()[:16]
    log(f"solve {solve_s * 1e3:.1f} ms: {result['outer_iterations']} outer / {result['inner_iterations']} inner, "
        f"solution sha256 {digest}")
    pass_bytes = 40 * n + 12 * e
    sweep = []
    for rows in (4, 8):
        for target in args.targets:
            batch.set_shape(rows, target)
            us = batch.time_pass(200)
            sweep.append(dict(rows=rows, target=target, us=us, algorithmic_gbs=pass_bytes / us / 1e3))
            log(f"pass rows={rows} target={target}: {us:.2f} us, {pass_bytes / us / 1e3:.0f} GB/s")
    stream = torch.cuda.Stream()
    apply = apply_roofline(operator, torch, stream, board=args.board)
    copy = copy_bandwidth(torch, stream, pass_bytes // 2 // 16 * 16)
    ceilings = stream_ceilings(torch, stream, n)
    best = min(sweep, key=lambda row: row["us"])
    return {
        "metric": f"fused Jacobi-PCG pass GB/s (algorithmic 40N+12E), {args.board} board {c}x{r}x{s}",
        "value": best["algorithmic_gbs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": 200,
        "warmup": 3,
        "ms_per_step": best["us"] / 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 sums)",
        "data": f"synthetic ({args.board} board generator)",
        "config": {"workload": f"{args.board} board {c}x{r}x{s} elements ({n:,} nodes)", "pass_sweep": sweep,
                   "pass_bytes": pass_bytes, "apply_us": apply["per_launch_us"], "apply_sweep": apply["sweep"],
                   "apply_algorithmic_gbs": apply["roofline"]["achieved"], "copy_gbs_at_pass_bytes": copy,
                   "stream_ceilings": ceilings,
                   "solve_ms": solve_s * 1e3, "inner_iterations": result["inner_iterations"],
                   "solution_sha256_16": digest,
                   "parallelism": "single GPU"},
    }


def _pass_bytes(op) -> int:
    """Algorithmic bytes of one fused PCG pass over one slot of ``op``."""

    s, r, c = op.element_shape
    return 40 * op.size + 12 * s * r * c


def serve_leg(args, torch, *, rates, duration: float, drain_limit: float) -> dict:
    """The serving stream: Poisson arrivals of thermal (70 %) and DC (30 %)
    requests of 1..16 right-hand sides on pool boards of the uneven-batch
    size distribution, admitted into a 200 GB device budget and iterated by
    continuous batching (tile shape tuned on the pool); against the same
    request mix solved one right-hand side at a time through the
    single-solve path."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from pcbhip import _capi, serving, workloads

    log("serve: building the pool")
    pool = workloads.serving_pool(seed=5, thermal_boards=args.pool_thermal, dc_boards=args.pool_dc)
    thermal = [b for b in pool if b.kind == "thermal"]
    dc = [b for b in pool if b.kind == "dc"]
    config = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000).to_c()
    rng = np.random.default_rng(5)

    def make_requests(rate: float) -> list:
        out = []
        for i, t in enumerate(serving.poisson_arrivals(rng, rate, duration)):
            group = thermal if (rng.random() < 0.7 or not dc) else dc
            board = group[int(rng.integers(len(group)))]
            m = int(rng.integers(1, 17))
            out.append(serving.ServeRequest(i, board.operator, board.rhs[:m], config, kind=board.kind,
                                            arrival_s=float(t)))
        return out

    # Single-solve path on the same mix: every pool board's first and last
    # right-hand side solved alone through a reusable one-slot batch; a
    # served request of m right-hand sides on board b then costs m times
    # board b's mean single-solve time on that path.
    single_s = {}
    for b in pool:
        sb = _capi.Batch([b.operator])
        sb.set_shape(args.rows, args.target)
        sb.set(0, b.rhs[0], None, config)
        sb.run()
        times = []
        for rhs in (b.rhs[0], b.rhs[-1]):
            sb.set(0, rhs, None, None)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sb.run()
            sb.result(0)
            times.append(time.perf_counter() - t0)
        single_s[id(b.operator)] = float(np.mean(times))
        sb.close()
    log(f"serve: single-solve path, {len(pool)} boards, mean {1e3 * np.mean(list(single_s.values())):.1f} ms per solve")
    cpu_sample = None
    if not args.no_cpu:
        from oracle.native import threads_default

        picks = thermal[:4] + dc[:2]
        cpu_sample = cpu_node_iteration_cost([(b.front_end, b.rhs[0]) for b in picks], threads_default())
        log(f"serve: cpu sample {cpu_sample['seconds']:.1f} s for {len(picks)} pool boards")
    lines = []
    shape = None
    for rate in rates:
        requests = make_requests(rate)
        server = serving.SolveServer(max_slots=args.max_slots, memory_budget_bytes=200e9,
                                     steps_per_replay=args.replay_steps, max_running_nodes=args.max_running_nodes,
                                     tune_with=[b.operator for b in pool])
        shape = server.shape
        log(f"serve: rate {rate}/s, {len(requests)} requests, {sum(len(r.rhs) for r in requests)} right-hand sides, "
            f"tile shape {shape}")
        elapsed = serving.run_arrivals(server, requests, log=log, time_limit_s=duration + drain_limit)
        done = server.completed
        solves = sum(len(r.results) for r in done)
        unfinished = len(requests) - len(done)
        by_id = {r.request_id: r for r in requests}
        single_time = sum(len(r.results) * single_s[id(by_id[r.request_id].operator)] for r in done)
        single_rate = solves / single_time if single_time > 0 else None
        summary = serving.latency_summary(done) if unfinished == 0 else {"p50_ms": None, "p99_ms": None}
        info = server.backend.info()
        # Algorithmic pass bytes of the inner iterations the served solves did,
        # over the device time of the server's replays.
        work = sum(x["inner_iterations"] * _pass_bytes(by_id[r.request_id].operator) for r in done for x in r.results)
        replay_s = info["replay_ms"] / 1e3
        cpu_line = None
        if cpu_sample is not None:
            node_its = sum(by_id[r.request_id].operator.size * max(x["inner_iterations"], 1)
                           for r in done for x in r.results)
            cpu_s = node_its * cpu_sample["seconds_per_node_iteration"]
            cpu_line = dict(cpu_sample, solves_per_s=solves / cpu_s,
                            sample_note="whole solves of 4 thermal and 2 DC pool boards on the host, scaled by the "
                                        "served solves' node-iterations")
        line = {
            "rate_per_s": rate,
            "requests": len(requests),
            "unfinished": unfinished,
            "offered_solves_per_s": sum(len(r.rhs) for r in requests) / duration,
            "solves_per_s": solves / elapsed,
            "single_solve_path_solves_per_s": single_rate,
            "speedup_vs_single_solve_path": (solves / elapsed) / single_rate if single_rate else None,
            "p50_ms": summary["p50_ms"],
            "p99_ms": summary["p99_ms"],
            "p50_ms_thermal": serving.latency_summary([r for r in done if r.kind == "thermal"])["p50_ms"],
            "p50_ms_dc": serving.latency_summary([r for r in done if r.kind == "dc"])["p50_ms"],
            "converged": int(sum(sum(x["converged"] for x in r.results) for r in done)),
            "inner_iterations_thermal": int(sum(x["inner_iterations"] for r in done if r.kind == "thermal"
                                                for x in r.results)),
            "inner_iterations_dc": int(sum(x["inner_iterations"] for r in done if r.kind == "dc" for x in r.results)),
            "pass_gbs_over_replays": work / replay_s / 1e9 if replay_s > 0 else None,
            "peak_requests_in_flight": server.peak_running_requests,
            "peak_slots": server.peak_running_slots,
            "peak_admitted_gb": server.peak_bytes / 1e9,
            "replays": info["replays"],
            "graph_builds": info["graph_builds"],
            "pumps": server.pumps,
            "pump_s": server.pump_s,
            "native_step_s": server.step_s,
            "blocked_on_device_s": info["wait_ms"] / 1e3,
            "device_replay_s": replay_s,
            "elapsed_s": elapsed,
            "cpu": cpu_line,
        }
        log("serve: " + json.dumps(line))
        lines.append(line)
        server.close()
    return {
        "workload": f"Poisson arrivals, 70 % thermal / 30 % DC, 1..16 right-hand sides, "
                    f"{len(thermal)} + {len(dc)} pool boards, {duration:g} s of arrivals per rate",
        "tile_shape": shape,
        "rates": lines,
    }


def run_serve(args, torch) -> dict:
    """The serving stream as its own line (``--workload serve``)."""

    leg = serve_leg(args, torch, rates=args.rates, duration=args.duration, drain_limit=args.drain_limit)
    best = max(leg["rates"], key=lambda l: l["solves_per_s"])
    return {
        "metric": "served MPIR solves/s at 1e-10 rel. residual with p50/p99 latency",
        "value": best["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded serving pool, seed 5)",
        "config": dict(leg, single_solve_path_solves_per_s=best["single_solve_path_solves_per_s"],
                       speedup_vs_single_solve_path=best["speedup_vs_single_solve_path"], parallelism="single GPU"),
    }


def run_coupling(args, torch) -> dict:
    """Electro-thermal loops as a served workload: the seeded scenarios run
    one after another (each loop's DC and thermal solves through the
    single-solve path, as the reference runs them) and then all at once, their
    solves continuously batched on one server (``run_electro_thermal_many``)."""

    from multiphysics.staggered_coupling import run_electro_thermal, run_electro_thermal_many
    from pcbhip import workloads

    scenarios = workloads.coupling_scenarios(count=args.coupling_count)
    log(f"coupling: {len(scenarios)} scenarios")
    for _ in range(args.warmup):
        run_electro_thermal(scenarios[0])
        run_electro_thermal_many(scenarios[:2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alone = [run_electro_thermal(s) for s in scenarios]
    sequential_s = time.perf_counter() - t0
    log(f"sequential: {sequential_s:.2f} s")
    served_s = []
    for step in range(args.steps):
        t0 = time.perf_counter()
        served = run_electro_thermal_many(
            scenarios, on_done=lambda i: log(f"served run {step}: loop {i} done at {time.perf_counter() - t0:.2f} s"))
        served_s.append(time.perf_counter() - t0)
    log(f"served: {np.mean(served_s):.2f} s per run")
    per_run = float(np.mean(served_s))
    solves = sum(2 * r.iterations for r in served)
    bitwise = all(np.array_equal(a.thermal.temperature_k, b.thermal.temperature_k, equal_nan=True)
                  and np.array_equal(a.electrical.potential_v, b.electrical.potential_v, equal_nan=True)
                  for a, b in zip(alone, served))
    return {
        "metric": "coupled electro-thermal loops/s (16 voltage-driven boards, served together)",
        "value": len(scenarios) / per_run,
        "unit": "loops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_run * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded coupling scenarios, seed 6)",
        "config": {
            "workload": f"{len(scenarios)} current loops, 8..48 x 32..256 elements, 2 copper layers + "
                        "3-slab thermal stack, sigma(T) with Aitken relaxation",
            "sequential_loops_per_s": len(scenarios) / sequential_s,
            "speedup_vs_one_loop_at_a_time": sequential_s / per_run,
            "device_solves": solves,
            "served_solves_per_s": solves / per_run,
            "converged": int(sum(r.converged for r in served)),
            "coupling_iterations": [r.iterations for r in served],
            "bitwise_equal_alone_and_served": bool(bitwise),
            "parallelism": "single GPU",
        },
    }


# The reference's published two-level MPIR solve times on its 4-slab fixture
# (docs/THERMAL_MPIR_FEM.md:472-482, 522-526; 1 W trace, MPIRConfig(
# max_outer_iterations=16), construction excluded): elements -> ms.
PUBLISHED_FIXTURE_MS = {
    50: {"xeon_8581c_16t": 26.6, "axion_c4a_32t": 23.0, "outer_inner": (7, 584)},
    100: {"xeon_8581c_16t": 94.0, "axion_c4a_32t": 74.0, "outer_inner": (7, 795)},
    200: {"xeon_8581c_16t": 239.7, "axion_c4a_32t": 218.4, "outer_inner": (6, 891)},
}


def _timed_ms(fn, repeats: int, warmups: int):
    """Median of ``repeats`` timed calls after ``warmups`` (the reference's
    _timed, experiments/thermal_native_benchmark.py:64-78)."""

    result = None
    for _ in range(warmups):
        result = fn()
    samples = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        result = fn()
        samples.append((time.perf_counter() - t0) * 1e3)
    return result, float(np.median(samples))


def run_two_level(args, torch) -> dict:
    """The two-level preconditioner against Jacobi: the reference's own
    4-slab fixture at its three published sizes (single solves through the
    front end's solve_mpir, construction timed apart), then the 512 x 512 x 16
    board."""

    from electrical.matrix_free_mpir_fem import MPIRConfig, solve_mpir
    from electrical.matrix_free_mpir_fem.solver import jacobi_default_config
    from pcbhip import workloads
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    cases = []
    fixture = {}
    for elements in (50, 100, 200):
        problem = workloads.four_slab_fixture(elements)
        row = {"elements": elements, "nodes": problem.mesh.size}
        for name, precond, config in (("two_level", "two-level", MPIRConfig(max_outer_iterations=16)),
                                      ("jacobi", "jacobi", jacobi_default_config(16))):
            t0 = time.perf_counter()
            operator = MatrixFreeThermalOperator(problem, preconditioner=precond)
            construction_ms = (time.perf_counter() - t0) * 1e3
            rhs = operator.build_rhs(operator.default_reference_temperature())
            result, ms = _timed_ms(lambda: solve_mpir(operator, rhs, config=config), repeats=5, warmups=1)
            row[name] = {"solve_ms": ms, "construction_ms": construction_ms, "converged": result.converged,
                         "outer": result.outer_iterations, "inner": result.inner_iterations,
                         "relative_residual": result.relative_residual}
            if precond == "two-level":
                row[name]["coarse_size"] = operator.coarse_correction.coarse_size
                row[name]["block"] = operator.coarse_correction.block
        published = PUBLISHED_FIXTURE_MS[elements]
        row["published_cpu_ms"] = published
        row["speedup_vs_xeon_16t"] = published["xeon_8581c_16t"] / row["two_level"]["solve_ms"]
        log(f"fixture {elements}: two-level {row['two_level']['solve_ms']:.2f} ms "
            f"({row['two_level']['outer']}/{row['two_level']['inner']}), jacobi {row['jacobi']['solve_ms']:.2f} ms "
            f"({row['jacobi']['outer']}/{row['jacobi']['inner']})")
        fixture[elements] = row
    if not args.no_board:
        problem = workloads.board_512x512x16(size=args.size, slabs=args.slabs)
        row = {"elements": f"{args.slabs}x{args.size}x{args.size}", "nodes": problem.mesh.size}
        for name, precond in (("two_level", "two-level"), ("jacobi", "jacobi")):
            t0 = time.perf_counter()
            operator = MatrixFreeThermalOperator(problem, preconditioner=precond)
            construction_ms = (time.perf_counter() - t0) * 1e3
            rhs = operator.build_rhs(operator.default_reference_temperature())
            config = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000)
            result, ms = _timed_ms(lambda: solve_mpir(operator, rhs, config=config), repeats=args.steps,
                                   warmups=args.warmup)
            row[name] = {"solve_ms": ms, "construction_ms": construction_ms, "converged": result.converged,
                         "outer": result.outer_iterations, "inner": result.inner_iterations,
                         "ms_per_inner_iteration": ms / max(1, result.inner_iterations)}
            log(f"board {name}: {ms:.1f} ms ({result.outer_iterations}/{result.inner_iterations}), "
                f"construction {construction_ms:.0f} ms")
        cases.append(row)
    top = fixture[200]
    value = 1e3 / top["two_level"]["solve_ms"]
    return {
        "metric": "two-level MPIR solves/s at 1e-10 on the reference's 4-slab fixture (202,005 nodes)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 5,
        "warmup": 1,
        "ms_per_step": top["two_level"]["solve_ms"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": value / (1e3 / PUBLISHED_FIXTURE_MS[200]["xeon_8581c_16t"]),
        "dtype": "f32 inner / f64 outer",
        "data": "the reference's deterministic 4-slab fixture",
        "config": {"workload": "4-slab fixture 200 x 200 elements, two-level, MPIRConfig(max_outer_iterations=16)",
                   "vs_baseline_against": "published C++ 16-thread Xeon Platinum 8581C solve, 239.7 ms",
                   "fixture": fixture, "boards": cases, "parallelism": "single GPU"},
    }


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--size", type=int, default=512)
    parser.add_argument("--slabs", type=int, default=16)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-oracle", action="store_true",
                        help="skip the FP64 oracle solve that states the GPU solution's error")
    parser.add_argument("--no-legs", action="store_true",
                        help="the board solve alone: no batch, multi-excitation or served summaries")
    parser.add_argument("--leg-duration", type=float, default=3.0,
                        help="seconds of arrivals of the served summary in the default line")
    parser.add_argument("--rows", type=int, default=4, help="rows per wave strip of the FP32 inner stencil")
    parser.add_argument("--target", type=int, default=2048, help="stencil workgroups per launch before z blocks split")
    parser.add_argument("--workload", default="board512",
                        choices=["board512", "batch200", "multirhs64", "serve", "apply_engines", "coupling",
                                 "two_level", "pass"])
    parser.add_argument("--board", default="standard", choices=["standard", "general"],
                        help="board512, pass: the benchmark board or the board with no stencil shortcuts")
    parser.add_argument("--targets", type=int, nargs="+", default=[512, 1024, 2048, 4096],
                        help="pass: stencil workgroup targets to sweep")
    parser.add_argument("--no-board", action="store_true", help="two_level: the fixture sizes only")
    parser.add_argument("--preconditioner", default="jacobi", choices=["jacobi", "two-level"],
                        help="batch200: the boards' preconditioner")
    parser.add_argument("--coupling-count", type=int, default=16, help="scenarios of the coupling workload")
    parser.add_argument("--rates", type=float, nargs="+", default=[50.0, 100.0, 200.0, 400.0])
    parser.add_argument("--duration", type=float, default=10.0)
    parser.add_argument("--drain-limit", type=float, default=60.0)
    parser.add_argument("--pool-thermal", type=int, default=16)
    parser.add_argument("--pool-dc", type=int, default=8)
    parser.add_argument("--max-slots", type=int, default=512)
    parser.add_argument("--replay-steps", type=int, default=16)
    parser.add_argument("--max-running-nodes", type=int, default=None)
    parser.add_argument("--dc-size", type=int, default=256)
    parser.add_argument("--block", type=int, default=1, help="right-hand sides per block-apply group")
    parser.add_argument("--engine", type=int, default=2, choices=[2, 3], help="bf16 terms of the matrix-core action")
    parser.add_argument("--count", type=int, default=200, help="boards in the batch200 workload")
    parser.add_argument("--compact-percent", type=int, default=25,
                        help="batch200: share of finished boards (percent) after which their tiles leave the "
                             "launches; 100 keeps every tile table for the whole run")
    args = parser.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU sharding is design-only"}))
        return 2

    import torch

    from pcbhip import _capi

    _capi.require_device(0)
    torch.cuda.init()
    runs = {"board512": run_board, "batch200": run_batch200, "multirhs64": run_multirhs64, "serve": run_serve,
            "apply_engines": run_apply_engines, "coupling": run_coupling, "two_level": run_two_level,
            "pass": run_pass}
    print(json.dumps(runs[args.workload](args, torch)), flush=True)
    return 0


def run_board(args, torch) -> dict:
    """The default line: one board's MPIR solve (the headline), the roofline
    of its dominant kernel, and -- unless ``--no-legs`` -- the batched,
    served and many-excitation workloads as summaries beside it, and the CPU
    baseline."""

    from electrical.matrix_free_mpir_fem import MPIRConfig
    from pcbhip import _capi
    from thermal.matrix_free_mpir_fem import MatrixFreeThermalOperator

    log("building board")
    problem = _board(args)
    s, r, c = problem.mesh.element_grid_shape
    nodes = problem.mesh.size
    if args.board == "general":
        workload = (f"general board {c}x{r}x{s} elements ({nodes:,} nodes): jittered per-element k, "
                    "films on both faces and the sides")
    else:
        workload = f"board {c}x{r}x{s} elements ({nodes:,} nodes), seeded copper/FR-4, 8 sources"
    operator = MatrixFreeThermalOperator(problem, preconditioner="jacobi")
    reference = operator.default_reference_temperature()
    rhs = operator.build_rhs(reference)
    config = MPIRConfig(relative_tolerance=1e-10, max_outer_iterations=60, max_inner_iterations=100000)
    batch = _capi.Batch([operator.device_operator])
    batch.set_shape(args.rows, args.target)
    batch.set(0, rhs, None, config.to_c())

    log("warmup")
    for _ in range(args.warmup):
        batch.run()
    latencies = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        batch.run()
        latencies.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    x, result = batch.result(0)
    log(f"solve: {result['outer_iterations']} outer / {result['inner_iterations']} inner, "
        f"rel {result['relative_residual']:.2e}, {elapsed / args.steps * 1e3:.1f} ms/solve")

    fused = pass_roofline(operator, batch, args.rows, args.target, board=args.board)
    log(f"pcg pass: {fused['per_launch_us']:.1f} us, {fused['roofline']['achieved']:.0f} GB/s")
    stream = torch.cuda.Stream()
    roof = apply_roofline(operator, torch, stream, board=args.board)
    log(f"apply: {roof['per_launch_us']:.1f} us, {roof['roofline']['achieved']:.0f} GB/s")
    copy_gbs = copy_bandwidth(torch, stream)
    log(f"copy: {copy_gbs:.0f} GB/s")
    # The same copy moving only as many bytes as one apply (or one pass) does:
    # the ramp-up and tail of a ~20 us launch cost what a 1 GiB copy hides.
    copy_at_apply = copy_bandwidth(torch, stream, roof["roofline"]["algorithmic_bytes"] // 2 // 16 * 16)
    copy_at_pass = copy_bandwidth(torch, stream, fused["roofline"]["algorithmic_bytes"] // 2 // 16 * 16)
    log(f"copy at apply bytes: {copy_at_apply:.0f} GB/s, at pass bytes: {copy_at_pass:.0f} GB/s")
    batch.close()
    line = {
        "metric": f"MPIR solves/s at 1e-10 rel. residual ({c}x{r}x{s} board)",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": f"synthetic (seeded {args.board} board generator)",
        "config": {
            "workload": workload,
            "solver": "MPIR, FP32 Jacobi-PCG inner (rtol 2e-3), FP64 outer to 1e-10",
            "converged": result["converged"],
            "relative_residual": result["relative_residual"],
            "outer_iterations": result["outer_iterations"],
            "inner_iterations": result["inner_iterations"],
            "latency_p50_ms": float(np.percentile(latencies, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(latencies, 99) * 1e3),
            "apply_gdof_s": roof["gdof_s"],
            "apply_us": roof["per_launch_us"],
            "apply_kernel": roof["kernel"],
            "apply_sweep": roof["sweep"],
            "apply_roofline": roof["roofline"],
            "pass_kernel": fused["kernel"],
            "pass_us": fused["per_launch_us"],
            "stream_copy_gbs": copy_gbs,
            "stream_copy_gbs_at_apply_bytes": copy_at_apply,
            "stream_copy_gbs_at_pass_bytes": copy_at_pass,
            "parallelism": "single GPU",
        },
        "roofline": fused["roofline"],
    }
    # The heat budget of the GPU solution (solve.py:255-280), on the
    # operator's own host and device actions.
    temperature = x.reshape(problem.mesh.node_shape) + reference
    residual = operator.unconstrained_residual(temperature, None)
    fixed_heat = -float(np.sum(residual[operator.fixed_nodes.reshape(-1)]))
    heat_input = float(np.sum(operator.nodal_load()))
    convective = float(np.sum(operator.convective_heat(temperature)))
    line["config"]["heat_input_w"] = heat_input
    line["config"]["heat_balance_error_w"] = heat_input - convective - fixed_heat
    if not args.no_cpu:
        # Every host figure below runs the port built for this host.
        from oracle import native

        native.use_native_build()
    if not args.no_legs:
        # The batched, multi-excitation and served workloads on short
        # schedules: their speed-ups over the single-solve path on the same
        # work, and the served latency at 50 requests/s.
        line["config"]["board_64x64x8_plain_fp64"] = plain_fp64_leg(args, torch)
        batched = batch_leg(args, torch, count=args.count, warmup=1, steps=1)
        many = multirhs_leg(args, torch, warmup=1, steps=1, variants=("slots", "block_mfma"))
        served = serve_leg(args, torch, rates=[50.0], duration=args.leg_duration, drain_limit=args.drain_limit)
        line["config"]["batch_200_uneven_boards"] = batched
        line["config"]["dc_board_64_excitations"] = many
        line["config"]["served_stream_50_per_s"] = dict(served["rates"][0], workload=served["workload"],
                                                       tile_shape=served["tile_shape"])
    if not args.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(problem, rhs, x, oracle_error=not args.no_oracle,
                                            calibrate=not args.no_legs)
    return line


if __name__ == "__main__":
    sys.exit(main())

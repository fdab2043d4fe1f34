#!/usr/bin/env python
"""Benchmark of the pressure-equation path on one MI355X.

Workload (one "step"): FP64 PCG with the reference's DIC preconditioner to a
relative residual of 1e-8 (tolerance 1e-300, the reference's stopping rule) on
the 128^3 unit-cube hex mesh: 2,097,152 cells, 6,242,304 internal faces,
vertices jittered by U[-0.2h, 0.2h] (boundary vertices tangentially), cells
randomly permuted (seed 0) then RCM-renumbered, face diffusivity
lognormal(0, 1), p = 0 on the x = 0 patch, zero-mean uniform right-hand side.
A step packs the LDU coefficients, factors DIC and runs the device PCG loop;
the assembled system is resident in HBM before timing starts.

Also measured in the same run:
  * the SpMV kernel (pfm_spmv) alone, HIP events on the launch stream, over
    four independent copies of the 128^3 system launched in turn (800 MB,
    beyond the 256 MB MALL, so every launch streams from HBM): algorithmic
    bytes 24N + 24F per launch -> GB/s vs the 8 TB/s HBM peak; beside it the
    same kernel on one MALL-resident copy, a plain streaming read of the same
    byte count (the achievable ceiling), and the HBM bytes per launch from two
    rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over a short run of this
    script (--roofline-probe), scaled by what the counters read for a copy of
    known size in the same run;
  * the adjoint: d(sum w*x)/d(gamma) with forward and adjoint solves to rel
    1e-10, as a ratio to the forward solve time;
  * the precision section: the 256^3 system (lognormal(0, 1.5) diffusivity)
    solved by FP64 pair/Galerkin GAMG and by its FP32-in-FP64-refinement form
    to rel 1e-8 (speed-up, true residuals, error against the FP64 solution),
    and a bounded PCG-DIC run on it (the reference algorithm stagnates there);
  * the CPU baseline: the reference's PCG-DIC op sequence in torch on the
    host cores (torch's default threads; oracle.torch_ref), a bounded number
    of iterations scaled to the device solve's iteration count, plus the DIC
    factor of the oracle's C restatement.

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per 1e-8 Poisson solve @2.1M cells, SpMV GB/s vs HBM peak, adjoint/fwd"
HBM_PEAK_GBS = 8000.0
ROOFLINE_COPIES = 4        # independent 128^3 systems the roofline launches rotate over
CALIB_BYTES = 256 << 20    # the counter-calibration copy of the roofline probe


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, d, lo, up, gpu_iters: int, budget_s: float):
    """The reference's PCG-DIC op sequence in torch on the host cores
    (oracle.torch_ref: gather + scatter_add products, torch.dot, torch.norm,
    r / d), torch's default intra-op threads; a bounded sample of iterations,
    scaled to gpu_iters, plus the DIC factor (the oracle's C restatement of
    the reference's sequential recurrence)."""
    import numpy as np
    import torch

    from oracle import precond as opre
    from oracle import torch_ref

    t0 = time.perf_counter()
    rD = opre.dic_build(d, up, case.owner, case.neighbour)
    t_fac = time.perf_counter() - t0
    A = torch_ref.LduCpu(d, lo, up, case.owner, case.neighbour)
    x0 = np.zeros(case.n_cells)
    # time a few iterations first, then size the sample to the budget
    t0 = time.perf_counter()
    torch_ref.pcg(A, case.b, x0, rD, tolerance=1e-300, rel_tol=1e-300, max_iter=3)
    t3 = (time.perf_counter() - t0) / 3
    k = int(max(3, min(gpu_iters, budget_s / max(t3, 1e-6))))
    t0 = time.perf_counter()
    torch_ref.pcg(A, case.b, x0, rD, tolerance=1e-300, rel_tol=1e-300, max_iter=k)
    t_k = time.perf_counter() - t0
    per_it = t_k / k
    ms = (t_fac + per_it * gpu_iters) * 1e3
    threads = torch.get_num_threads()
    return {"value": round(ms, 1), "unit": "ms per solve", "cores": threads, "kind": "port",
            "cpu_count": os.cpu_count(), "torch_threads": threads, "cpu_model": _cpu_model(),
            "sample": f"the reference's PCG-DIC op sequence in torch on the host (oracle.torch_ref, "
                      f"{threads} torch threads, os.cpu_count() = {os.cpu_count()}) on the same "
                      f"128^3 system: {k} iterations ({t_k:.1f} s, {per_it * 1e3:.1f} ms/iter) "
                      f"scaled to the GPU solve's {gpu_iters} iterations, plus the DIC factor "
                      f"(oracle C restatement, {t_fac * 1e3:.0f} ms)"}


def _roofline_copies(case, mat, dev, k: int):
    """k independent copies of the system (own topology handle, coefficients,
    packed values, x, y): launched in turn, no copy stays MALL-resident."""
    import numpy as np
    import torch

    from pyopenfoam_amd import LduMatrix

    out = []
    for i in range(k):
        m = LduMatrix(case.n_cells, torch.from_numpy(case.owner.copy()),
                      torch.from_numpy(case.neighbour.copy()), device=dev)
        m.diag, m.lower, m.upper = mat.diag.clone(), mat.lower.clone(), mat.upper.clone()
        x = torch.from_numpy(np.random.default_rng(100 + i).standard_normal(case.n_cells)).to(dev)
        out.append((m, m.packed(False), x, torch.empty_like(x)))
    return out


def _spmv_launcher(cps, dev):
    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.core.device import current_stream

    L, s = _lib.lib(), current_stream(dev)
    args = [(m.addressing.handle, _lib.ptr(m.diag), _lib.ptr(pk), _lib.ptr(x), _lib.ptr(y))
            for (m, pk, x, y) in cps]

    def launch(i):
        h, d, p, x, y = args[i % len(args)]
        L.pfm_spmv(h, d, p, x, y, s)
    return launch


def roofline_probe(args) -> int:
    """--roofline-probe: the short run the counter passes profile: rotating
    k_spmv launches over the roofline copies, then a device copy of
    CALIB_BYTES (pfm_bandwidth_probe mode 1) whose bytes are known."""
    import torch

    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.cases import poisson_case
    from pyopenfoam_amd.core.device import current_stream
    from pyopenfoam_amd.discretisation.fvm import assemble
    from pyopenfoam_amd.mesh.fv_mesh import FvMesh

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case = poisson_case(args.n, jitter=args.jitter, sigma=args.sigma, seed=0)
    mat = assemble(case, FvMesh.from_case(case, dev))
    cps = _roofline_copies(case, mat, dev, ROOFLINE_COPIES)
    launch = _spmv_launcher(cps, dev)
    for i in range(8 * ROOFLINE_COPIES):
        launch(i)
    a = torch.ones(CALIB_BYTES // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    L = _lib.lib()
    for _ in range(4):
        _lib.check(L.pfm_bandwidth_probe(1, _lib.ptr(a), _lib.ptr(b), CALIB_BYTES,
                                         current_stream(dev)), "copy probe")
    torch.cuda.synchronize()
    print("roofline probe done", file=sys.stderr, flush=True)
    return 0


def spmv_kernel_name(case) -> str:
    """The kernel pfm_spmv streams with: the paired 16-bit column offsets when
    every coupling's column is within 32767 of its row, or within 65535 on a
    mesh whose owners all precede their neighbours (the handle then keeps
    them), else the int32 columns."""
    import numpy as np

    nf = case.mesh.n_internal_faces
    d = case.neighbour.astype(np.int64) - case.owner[:nf]
    band = int(np.abs(d).max()) if nf else 0
    oriented = bool(nf == 0 or d.min() > 0)
    if (band <= 32767 or (oriented and band <= 65535)) and os.environ.get("PFM_SPMV_COLS") != "int32":
        return "k_spmv_p2<double, true>"
    return "k_spmv<double, true>"


def measure_traffic(args, timeout_s: float = 300.0):
    """HBM bytes per k_spmv launch: two rocprofv3 counter passes (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch) over --roofline-probe, each scaled by
    what the counter reads for the calibration copy of CALIB_BYTES in the same
    run.  Returns a dict, or None with the reason logged."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        log("traffic: rocprofv3 not found")
        return None
    res = {"command": []}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix=f"pfm_{counter.lower()}_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "-d", out, "-o", "pmc", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--roofline-probe", "--n", str(args.n),
               "--jitter", str(args.jitter), "--sigma", str(args.sigma)]
        res["command"].append(" ".join(cmd[:9] + ["--", "python", "bench.py"] + cmd[12:]))
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, env=env,
                               cwd="/tmp")
        except subprocess.TimeoutExpired:
            log(f"traffic: rocprofv3 {counter} pass timed out")
            return None
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            log(f"traffic: rocprofv3 {counter} pass wrote no counters (rc {r.returncode}): "
                f"{r.stderr[-400:]}")
            return None
        spmv, calib = [], []
        for fn in files:
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name, v = row.get("Kernel_Name", ""), float(row.get("Counter_Value", "nan"))
                    if "k_spmv<double, true>" in name or "k_spmv_p2<double, true>" in name:
                        spmv.append(v)
                    elif "k_bw_copy" in name:
                        calib.append(v)
        shutil.rmtree(out, ignore_errors=True)
        if not spmv or not calib:
            log(f"traffic: {counter} pass found {len(spmv)} k_spmv and {len(calib)} copy dispatches")
            return None
        kib = sum(spmv) / len(spmv)
        scale = (CALIB_BYTES / 1024) / (sum(calib) / len(calib))
        res[counter] = {"kib_per_dispatch_raw": round(kib, 1), "dispatches": len(spmv),
                        "scale_from_calibration_copy": round(scale, 4),
                        "bytes_per_launch": round(kib * 1024 * scale)}
        log(f"traffic {counter}: {kib:.1f} KiB raw x {scale:.3f} per k_spmv launch "
            f"({len(spmv)} dispatches)")
    res["bytes_per_launch"] = res["FETCH_SIZE"]["bytes_per_launch"] + \
        res["WRITE_SIZE"]["bytes_per_launch"]
    return res


def bench_roofline(case, mat, dev, args):
    """k_spmv from HBM: HIP events on the launch stream around rotating
    launches over ROOFLINE_COPIES independent copies; the same kernel on one
    MALL-resident copy; a plain streaming read of the same bytes."""
    import numpy as np
    import torch

    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.core.device import current_stream

    N, F = case.n_cells, case.mesh.n_internal_faces
    nbytes = 24 * N + 24 * F
    cps = _roofline_copies(case, mat, dev, ROOFLINE_COPIES)
    launch = _spmv_launcher(cps, dev)
    ref = mat.Ax_packed(cps[0][2])
    launch(0)
    torch.cuda.synchronize()
    assert torch.equal(cps[0][3], ref), "roofline copy differs from the solver's matrix"

    def timed(fn, reps):
        for i in range(2 * ROOFLINE_COPIES):
            fn(i)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(reps):
            fn(i)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    us_hbm = timed(launch, 50 * ROOFLINE_COPIES)
    us_res = timed(lambda i: launch(0), 200)
    L, s = _lib.lib(), current_stream(dev)
    nb16 = nbytes // 16 * 16
    bufs = [torch.ones(nb16 // 8, dtype=torch.float64, device=dev) for _ in range(ROOFLINE_COPIES)]
    part = torch.empty(int(L.pfm_bandwidth_probe_out_bytes()) // 8, dtype=torch.float64, device=dev)
    us_read = timed(lambda i: L.pfm_bandwidth_probe(0, _lib.ptr(bufs[i % len(bufs)]),
                                                    _lib.ptr(part), nb16, s),
                    20 * ROOFLINE_COPIES)
    del bufs, cps
    torch.cuda.empty_cache()
    gbs = lambda us, b=nbytes: b / (us * 1e-6) / 1e9
    log(f"SpMV from HBM ({ROOFLINE_COPIES} copies rotated): {us_hbm:.2f} us/launch, "
        f"{gbs(us_hbm):.0f} GB/s ({gbs(us_hbm) / HBM_PEAK_GBS:.3f} of 8 TB/s); MALL-resident "
        f"{us_res:.2f} us ({gbs(us_res):.0f} GB/s); streaming read of the same bytes "
        f"{gbs(us_read, nb16):.0f} GB/s")
    traffic = None if args.no_traffic else measure_traffic(args)
    return {"us_hbm": us_hbm, "us_resident": us_res, "gbs_hbm": gbs(us_hbm),
            "gbs_resident": gbs(us_res), "read_ceiling_gbs": gbs(us_read, nb16),
            "bytes": nbytes, "traffic": traffic}


def bench_batch(args, dev):
    """512 (default) jittered systems, side n ~ U{8..48}: batched PCG-DIC to
    rel 1e-8 (graph-replayed), the batched adjoint of J = sum(w*x), and an
    estimate of the same systems solved one at a time with the single-system
    device PCG (a size-stratified sample, scaled to the whole batch)."""
    import numpy as np
    import torch

    from pyopenfoam_amd import LduMatrix, PCGSolver
    from pyopenfoam_amd.cases import poisson_batch
    from pyopenfoam_amd.discretisation.fvm import LaplacianAssemble
    from pyopenfoam_amd.differentiable.linear_solver import _Ctx
    from pyopenfoam_amd.solvers.batch import BatchedLdu, BatchedPCGSolver, batch_poisson_solve

    t0 = time.perf_counter()
    bc = poisson_batch(args.batch_systems, args.batch_nmin, args.batch_nmax, seed=0,
                       jitter=args.jitter, sigma=args.sigma, workers=min(16, os.cpu_count() or 1))
    t_gen = time.perf_counter() - t0
    ns, N, F = bc.n_systems, bc.n_cells, int(bc.face_offsets[-1])
    log(f"batch: {ns} systems, {N} cells, {F} faces, built in {t_gen:.1f} s")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    t0 = time.perf_counter()
    B = BatchedLdu(bc.cell_offsets, bc.face_offsets, bc.owner_local, bc.neighbour_local, dev,
                   global_owner=bc.owner, global_neighbour=bc.neighbour)
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    log(f"batch handle + autotune {t_create:.1f} s: tile {B.tile_slices} slices, {B.n_tiles} tiles, "
        f"us/iter {B.autotune_us}")
    geo = (t(bc.mag), t(bc.delta), t(bc.vol), t(bc.dir_cells.astype(np.int32)), t(bc.dir_mag),
           t(bc.dir_delta))
    gam, gamb, b = t(bc.gamma), t(bc.dir_gamma), t(bc.b)
    d, lo, up = LaplacianAssemble.forward(_Ctx(), gam, gamb, B.global_addressing, *geo)
    max_iter = args.batch_max_iter
    solver = BatchedPCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=max_iter)
    x, res = solver.solve(B, d, lo, up, b)  # warm
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        x, res = solver.solve(B, d, lo, up, b)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    # true relative residual per system
    own, nei = B.global_addressing.owner32.long(), B.global_addressing.neighbour32.long()
    ax = d * x
    ax.index_add_(0, own, lo * x[nei])
    ax.index_add_(0, nei, up * x[own])
    seg = torch.repeat_interleave(torch.arange(ns, device=dev),
                                  t(np.diff(bc.cell_offsets)).long())
    r2 = torch.zeros(ns, dtype=torch.float64, device=dev).index_add_(0, seg, (b - ax) ** 2)
    b2 = torch.zeros(ns, dtype=torch.float64, device=dev).index_add_(0, seg, b * b)
    true_rel = (r2.sqrt() / b2.sqrt()).max().item()
    it = res.iterations
    rel_k = (r2.sqrt() / b2.sqrt()).cpu().numpy()
    bad = np.flatnonzero(~res.converged)
    log(f"batch solve: {ms:.1f} ms for {ns} systems, iterations max {it.max()} mean {it.mean():.0f}, "
        f"converged {ns - len(bad)}/{ns}, max true rel residual {true_rel:.2e}")
    for k in bad[:20]:
        log(f"  not converged: system {k} n={bc.sizes[k]} iterations {it[k]} status {res.status[k]} "
            f"final {res.final_residual[k]:.3e} initial {res.initial_residual[k]:.3e} "
            f"true rel {rel_k[k]:.3e}")
    conv = res.converged
    p50, p90, p99 = (np.percentile(it[conv], q) for q in (50, 90, 99)) if conv.any() else (0, 0, 0)
    log(f"  iterations of converged systems: p50 {p50:.0f} p90 {p90:.0f} p99 {p99:.0f} "
        f"max {it[conv].max() if conv.any() else 0}")
    # adjoint: J = sum over systems of sum(w*x)
    g = gam.clone().requires_grad_(True)
    gb = gamb.clone().requires_grad_(True)
    w = t(np.random.default_rng(1).standard_normal(N))
    asolver = BatchedPCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=max_iter)
    tf = tb = 0.0
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs = batch_poisson_solve(g, gb, b, asolver, B, geo)
        J = (w * xs).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        tf, tb = t1 - t0, time.perf_counter() - t1
        g.grad = None
        gb.grad = None
    log(f"batch adjoint: forward {tf * 1e3:.1f} ms, backward {tb * 1e3:.1f} ms, ratio {tb / tf:.3f}")
    # the same systems one at a time (size-stratified sample)
    order = np.argsort(bc.sizes, kind="stable")
    k_s = max(1, min(args.batch_seq_sample, ns))
    sample = order[np.linspace(0, ns - 1, k_s).round().astype(int)]
    ssolver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=max_iter)
    seq_s = 0.0
    for k in sample:
        c0, c1 = int(bc.cell_offsets[k]), int(bc.cell_offsets[k + 1])
        f0, f1 = int(bc.face_offsets[k]), int(bc.face_offsets[k + 1])
        m = LduMatrix(c1 - c0, torch.from_numpy(bc.owner_local[f0:f1]),
                      torch.from_numpy(bc.neighbour_local[f0:f1]), device=dev)
        m.diag, m.lower, m.upper = d[c0:c1], lo[f0:f1], up[f0:f1]
        _ = m.addressing
        xk = torch.zeros(c1 - c0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ssolver(m, b[c0:c1], xk)
        torch.cuda.synchronize()
        seq_s += time.perf_counter() - t0
    seq_ms = seq_s * 1e3 * ns / k_s
    log(f"one at a time: {seq_s * 1e3:.1f} ms for {k_s} sampled systems -> ~{seq_ms:.0f} ms "
        f"for {ns} ({seq_ms / ms:.1f}x the batched solve)")
    # algorithmic bytes: each system's PCG iteration moves 24 F_k + 112 N_k (the
    # single-system figure), counted over the iterations each system ran
    nk = np.diff(bc.cell_offsets).astype(np.float64)
    fk = np.diff(bc.face_offsets).astype(np.float64)
    per_iter = 24.0 * fk + 112.0 * nk
    batch_bytes = float((it.astype(np.float64) * per_iter).sum())
    gbs = batch_bytes / (ms * 1e-3) / 1e9
    full_us = float(B.autotune_us[B.tile_slices])
    full_gbs = float(per_iter.sum()) / (full_us * 1e-6) / 1e9
    log(f"batch bandwidth: {batch_bytes / 1e9:.1f} GB algorithmic over the solve, {gbs:.0f} GB/s "
        f"({gbs / HBM_PEAK_GBS:.3f} of peak); every system active: {full_gbs:.0f} GB/s "
        f"({full_gbs / HBM_PEAK_GBS:.3f})")
    return {"n_systems": ns, "n_cells": N, "n_internal_faces": F,
            "sizes": f"n ~ U{{{args.batch_nmin}..{args.batch_nmax}}} seed 0",
            "ms": round(ms, 2), "ms_per_system": round(ms / ns, 4), "max_iter": max_iter,
            "iterations_max": int(it.max()), "iterations_mean": round(float(it.mean()), 1),
            "all_converged": bool(res.converged.all()), "n_converged": int(conv.sum()),
            "not_converged": [{"system": int(k), "n": int(bc.sizes[k]), "iterations": int(it[k]),
                               "status": int(res.status[k]),
                               "true_rel_residual": float(rel_k[k])} for k in bad[:20]],
            "max_true_rel_residual": true_rel,
            "tile_slices": B.tile_slices, "n_tiles": B.n_tiles,
            "autotune_us_per_iter": {str(k): round(v, 2) for k, v in B.autotune_us.items()},
            "adjoint": {"forward_ms": round(tf * 1e3, 2), "backward_ms": round(tb * 1e3, 2),
                        "ratio": round(tb / tf, 3)},
            "algorithmic_bytes": batch_bytes, "gbs_algorithmic": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "full_load": {"us_per_iteration": round(full_us, 2),
                          "bytes_per_iteration": float(per_iter.sum()),
                          "gbs": round(full_gbs, 1),
                          "frac_of_hbm_peak": round(full_gbs / HBM_PEAK_GBS, 4)},
            "one_at_a_time_ms_est": round(seq_ms, 1), "one_at_a_time_sample": int(k_s),
            "speedup_vs_one_at_a_time": round(seq_ms / ms, 2),
            "setup_s": {"cases_host": round(t_gen, 1), "handle_autotune": round(t_create, 1)}}


def bench_geometry(case, dev, no_cpu: bool):
    """Device mesh geometry (pfm_mesh_geometry) on the case's mesh: kernel
    time from HIP events on the launch stream (5 launches), the whole call
    with host CSR and transfers, and the oracle's vectorised host geometry on
    a 40^3 sample scaled per cell."""
    import numpy as np
    import torch

    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.core.device import current_stream
    from pyopenfoam_amd.mesh.geometry import cell_faces, compute_geometry, face_csr

    m = case.mesh
    t0 = time.perf_counter()
    compute_geometry(m, device=dev)
    torch.cuda.synchronize()
    t_call = time.perf_counter() - t0
    fptr, fv = face_csr(m.faces)
    cptr, cf, mx = cell_faces(m.n_cells, m.owner, m.neighbour, fptr)
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)
    pts = up(m.points, torch.float64)
    d = [up(a, torch.int32) for a in (fptr, fv, m.owner, m.neighbour, cptr, cf)]
    nf, nfi, nc = m.n_faces, m.n_internal_faces, m.n_cells
    f64 = dict(dtype=torch.float64, device=dev)
    outs = [torch.empty(nf, 3, **f64), torch.empty(nf, 3, **f64), torch.empty(nf, **f64),
            torch.empty(nc, **f64), torch.empty(nc, 3, **f64), torch.empty(nf, **f64)]
    p = _lib.ptr
    L = _lib.lib()
    s = current_stream(dev)
    call = lambda: _lib.check(L.pfm_mesh_geometry(
        int(m.points.shape[0]), p(pts), nf, p(d[0]), p(d[1]), nfi, p(d[2]), p(d[3]), nc, p(d[4]),
        p(d[5]), mx, *[p(o) for o in outs], s), "pfm_mesh_geometry")
    call()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        call()
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / 5
    out = {"n_cells": nc, "n_faces": nf, "kernels_ms": round(k_ms, 3),
           "call_with_transfers_s": round(t_call, 3)}
    log(f"geometry: kernels {k_ms:.2f} ms, whole call {t_call:.2f} s")
    if not no_cpu:
        from oracle import mesh as omesh
        from pyopenfoam_amd.mesh.hexmesh import relabel_cells, structured_hex
        ms = structured_hex(40)
        ms = relabel_cells(ms, np.random.default_rng(0).permutation(ms.n_cells))
        t0 = time.perf_counter()
        omesh.geometry_vec(ms.points, ms.faces, ms.owner, ms.neighbour, ms.n_cells)
        t_o = time.perf_counter() - t0
        out["cpu_oracle"] = {"sample": "oracle geometry_vec (numpy, vectorised) on a 40^3 mesh",
                             "s": round(t_o, 2), "s_scaled_to_case": round(t_o * nc / ms.n_cells, 1)}
        log(f"geometry oracle 40^3: {t_o:.2f} s")
    return out


def bench_pressure(case, mesh, dev):
    """The pressure-equation step around the solve at the case's size: HIP
    event times (5 launches each) and GB/s on algorithmic bytes (every input
    read once, every output written once):
      assemble      56 B/internal face (own, nei, w, |S|, delta, phi; lower,
                    upper) + 32 B/cell (A_p, V; diag, source) + 12 B/boundary face
      face flux     48 B/face + 16 B/cell
      velocity      40 B/face (own, nei, w, S) + 72 B/cell (p, A_p, V, HbyA, U)
                    + 28 B/boundary face"""
    import numpy as np
    import torch

    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.core.device import current_stream
    from pyopenfoam_amd.solvers import (assemble_pressure_equation, correct_face_flux,
                                        correct_velocity)

    n, nfi, nfa = mesh.n_cells, mesh.n_internal_faces, mesh.n_faces
    nb = nfa - nfi
    rng = np.random.default_rng(11)
    t = lambda a: torch.from_numpy(a).to(dev)
    ap, phi, p = t(rng.uniform(1, 2, n)), t(rng.normal(size=nfa)), t(rng.normal(size=n))
    hb = t(rng.normal(size=(n, 3)))
    mesh.face_weights  # built once, outside the timings
    mat = assemble_pressure_equation(phi, ap, mesh)
    L, q = _lib.lib(), _lib.ptr
    h = mesh.ldu_addressing().handle
    w = mesh.face_weights[:nfi].contiguous()
    mag = mesh.face_areas_magnitude[:nfi].contiguous()
    dl = mesh.delta_coefficients[:nfi].contiguous()
    sf = mesh.face_areas.contiguous()
    vol = mesh.cell_volumes
    bptr, bface = mesh.boundary_csr
    s = current_stream(dev)
    out_f = torch.empty_like(phi)
    out_u = torch.empty_like(hb)
    calls = {
        "assemble": (lambda: L.pfm_pressure_assemble(
            h, nfa, q(bptr), q(bface), q(phi), q(ap), q(w), q(mag), q(dl), q(vol), q(mat.diag),
            q(mat.lower), q(mat.upper), q(mat.source), s), 56 * nfi + 32 * n + 12 * nb),
        "correct_face_flux": (lambda: L.pfm_correct_face_flux(
            h, nfa, q(phi), q(p), q(ap), q(w), q(mag), q(dl), q(out_f), s), 48 * nfi + 16 * n),
        "correct_velocity": (lambda: L.pfm_correct_velocity(
            h, nfa, q(bptr), q(bface), q(hb), q(p), q(ap), q(w), q(sf), q(vol), 0.7, q(out_u), s),
            40 * nfi + 72 * n + 28 * nb),
    }
    out = {}
    for name, (fn, nbytes) in calls.items():
        _lib.check(fn(), name)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / 5 * 1e3
        out[name] = {"us": round(us, 2), "algorithmic_bytes": int(nbytes),
                     "gbs": round(nbytes / (us * 1e-6) / 1e9, 1)}
        log(f"pressure step {name}: {us:.1f} us, {nbytes / (us * 1e-6) / 1e9:.0f} GB/s")
    return out


def bench_family(mat, dev):
    """The rest of the solver family on the benchmark matrix: DILU / DIC factor
    times (HIP events, 5 factors), fixed iteration counts of PBiCGStab (DILU),
    PBiCG (DILU) and SmoothSolver (Gauss-Seidel, one sweep per iteration),
    reported per iteration (tolerances 0; the reference's DILU PBiCGStab
    breaks down on this system after a few dozen iterations, see the status)."""
    import torch

    from pyopenfoam_amd import PBiCGSTABSolver
    from pyopenfoam_amd.solvers import (DICPreconditioner, DILUPreconditioner, PBiCGSolver,
                                        SmoothSolver)

    out = {}
    for name, cls in (("dilu_factor_ms", DILUPreconditioner), ("dic_factor_ms", DICPreconditioner)):
        cls(mat)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            cls(mat)
        e1.record()
        torch.cuda.synchronize()
        out[name] = round(e0.elapsed_time(e1) / 5, 3)
    x0 = torch.zeros(mat.n_cells, dtype=torch.float64, device=dev)
    runs = (("pbicgstab_dilu", lambda k: PBiCGSTABSolver(tolerance=0.0, rel_tol=0.0, max_iter=k,
                                                         preconditioner="DILU"), 400),
            ("pbicg_dilu", lambda k: PBiCGSolver(tolerance=0.0, rel_tol=0.0, max_iter=k,
                                                 preconditioner="DILU"), 400),
            ("smooth_gauss_seidel", lambda k: SmoothSolver(tolerance=0.0, rel_tol=0.0, max_iter=k,
                                                           smoother="GaussSeidel", n_sweeps=1), 20))
    for name, make, k in runs:
        mat.solve(make(5), x0, 0.0, 5)  # warm
        torch.cuda.synchronize()
        solver = make(k)
        t0 = time.perf_counter()
        mat.solve(solver, x0, 0.0, k)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        it = max(1, solver.last_info.iterations - 1)
        # the DILU solvers build their factor in the call: per iteration without it
        per = (ms - (out["dilu_factor_ms"] if "dilu" in name else 0.0)) / it
        out[name] = {"ms": round(ms, 2), "iterations": it, "ms_per_iteration": round(per, 4),
                     "status": int(getattr(solver.last_info, "status", 0))}
        log(f"family {name}: {ms:.1f} ms for {it} iterations ({per * 1e3:.1f} us each after "
            f"the factor), status {out[name]['status']}")
    log(f"family factors: DILU {out['dilu_factor_ms']} ms, DIC {out['dic_factor_ms']} ms")
    return out


def gs_launches_per_cycle(hier, solver) -> int:
    """Gauss-Seidel level launches of one V-cycle: levels above 1024 cells
    run one launch per dependency level of a symmetric sweep (forward and
    backward), n_pre + n_post sweeps per level (the smaller levels run inside
    the one-workgroup tail)."""
    import ctypes as C

    import numpy as np

    from pyopenfoam_amd import _lib

    L = _lib.lib()
    sweeps = solver._n_pre_smooth + solver._n_post_smooth
    total = 0
    for k, (n, nf) in enumerate(hier.level_sizes()):
        if n <= 1024:
            continue
        own, nei = hier.level_topology(k)
        o, q = _lib.host_i32(own), _lib.host_i32(nei)
        fl, bl = np.empty(n, np.int32), np.empty(n, np.int32)
        nfl, nbl = C.c_int32(), C.c_int32()
        _lib.check(L.pfm_gs_levels(n, int(o.shape[0]), _lib.ptr(o), _lib.ptr(q), _lib.ptr(fl),
                                   _lib.ptr(bl), C.byref(nfl), C.byref(nbl)), "pfm_gs_levels")
        total += sweeps * (nfl.value + nbl.value)
    return total


def bench_precision(args, dev):
    """The 256^3 jittered mesh with lognormal(0, 1.5) diffusivity.
    FP64 pair/Galerkin GAMG and its mixed-precision form (FP32 V-cycles on the
    levels above 4096 cells inside FP64 refinement on the true residual, FP64
    coarse levels and coarsest solve) to rel 1e-8; the SpMV at this size (1.6 GB
    per product, HBM-streaming); and PCG-DIC (the reference algorithm) for a
    bounded number of iterations, where it stagnates far above 1e-8."""
    import torch

    from pyopenfoam_amd import GAMGSolver, PCGSolver, _lib
    from pyopenfoam_amd.cases import poisson_case
    from pyopenfoam_amd.core.device import current_stream
    from pyopenfoam_amd.discretisation.fvm import assemble
    from pyopenfoam_amd.mesh.fv_mesh import FvMesh

    n, sigma = args.precision_n, args.precision_sigma
    t0 = time.perf_counter()
    case = poisson_case(n, jitter=args.jitter, sigma=sigma, seed=0)
    mesh = FvMesh.from_case(case, dev)
    mat = assemble(case, mesh)
    torch.cuda.synchronize()
    N, F = case.n_cells, case.mesh.n_internal_faces
    log(f"precision: {n}^3, {N} cells, {F} faces, lognormal(0,{sigma}), built in "
        f"{time.perf_counter() - t0:.1f} s")
    x0 = torch.zeros(N, dtype=torch.float64, device=dev)
    out = {"workload": f"{n}^3 hex mesh jitter {args.jitter}, permuted+RCM, lognormal(0,{sigma}) "
                       f"diffusivity, Dirichlet x0 patch, rel 1e-8 (tolerance 1e-300)",
           "n_cells": N, "n_internal_faces": F}
    # SpMV at this size
    pk = mat.packed(False)
    y = torch.empty_like(x0)
    xr = torch.ones_like(x0)
    L, s = _lib.lib(), current_stream(dev)
    h = mat.addressing.handle
    for _ in range(3):
        L.pfm_spmv(h, _lib.ptr(mat.diag), _lib.ptr(pk), _lib.ptr(xr), _lib.ptr(y), s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        L.pfm_spmv(h, _lib.ptr(mat.diag), _lib.ptr(pk), _lib.ptr(xr), _lib.ptr(y), s)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 20 * 1e3
    sb = 24 * N + 24 * F
    out["spmv"] = {"us": round(us, 2), "algorithmic_bytes": sb,
                   "gbs": round(sb / (us * 1e-6) / 1e9, 1),
                   "frac": round(sb / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"precision: SpMV {us:.1f} us, {sb / (us * 1e-6) / 1e9:.0f} GB/s")
    bn = mat.source.norm().item()
    kw = dict(agglomeration="pair", coarse="galerkin", max_levels=20)
    x64 = None
    for name, extra in (("gamg_fp64", {}),
                        ("gamg_mixed", dict(precision="mixed", inner_rel_tol=0.1,
                                            max_outer=100))):
        gs = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=args.precision_max_cycles,
                        **kw, **extra)
        mat.solve(gs, x0, 1e-300, 2)  # warm: hierarchy, coarse handles, workspace
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xg, it, _ = mat.solve(gs, x0, 1e-300, args.precision_max_cycles)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        info = gs.last_info
        rt = (mat.source - mat.Ax(xg)).norm().item() / bn
        cyc = info.inner_iterations if extra else int(it)
        rec = {"ms": round(ms, 1), "cycles": int(cyc), "converged": bool(info.converged),
               "true_rel_residual": rt, "levels": gs.hierarchy(mat).n_levels}
        if extra:
            rec["refinement_steps"] = int(it) - 1
            rec["inner_rel_tol"] = 0.1
            if x64 is not None:
                rec["speedup_vs_fp64"] = round(out["gamg_fp64"]["ms"] / ms, 3)
                rec["rel_err_vs_fp64_solution"] = ((xg - x64).norm() / x64.norm()).item()
        else:
            x64 = xg
        out[name] = rec
        log(f"precision {name}: {ms:.0f} ms, {cyc} cycles, converged {info.converged}, "
            f"true rel residual {rt:.2e}" + (f", {rec.get('speedup_vs_fp64')}x FP64, error "
                                              f"{rec.get('rel_err_vs_fp64_solution', 0):.2e}"
                                              if extra else ""))
    # the reference algorithm on this system: bounded PCG-DIC, FP64 and mixed
    for name, extra in (("pcg_dic_fp64", {}), ("pcg_dic_mixed", dict(precision="mixed"))):
        k = args.precision_pcg_iters
        ps = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=k, **extra)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xp, it, _ = mat.solve(ps, x0, 1e-300, k)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        rt = (mat.source - mat.Ax(xp)).norm().item() / bn
        inner = getattr(ps.last_info, "inner_iterations", None)
        iters = int(inner) if extra else int(it) - 1
        out[name] = {"iterations": iters, "ms": round(ms, 1),
                     "us_per_iteration": round(ms * 1e3 / max(iters, 1), 2),
                     "converged": bool(ps.last_info.converged), "true_rel_residual": rt}
        log(f"precision {name}: {iters} iterations in {ms:.0f} ms, converged "
            f"{ps.last_info.converged}, true rel residual {rt:.2e}")
    del mat, mesh, case
    torch.cuda.empty_cache()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--max-iter", type=int, default=200000)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--inner-rel-tol", type=float, default=1e-3)
    ap.add_argument("--jitter", type=float, default=0.2)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--history", default=None, help="write the residual history (.npy) here")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-gamg", action="store_true")
    ap.add_argument("--no-pressure", action="store_true")
    ap.add_argument("--no-family", action="store_true")
    ap.add_argument("--gamg-max-iter", type=int, default=1000)
    ap.add_argument("--gamg-variants", default="",
                    help="comma-separated GAMG variants to run (default: all)")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-nmin", type=int, default=8)
    ap.add_argument("--batch-nmax", type=int, default=48)
    ap.add_argument("--batch-max-iter", type=int, default=30000)
    ap.add_argument("--batch-seq-sample", type=int, default=32,
                    help="systems solved one at a time to estimate the unbatched loop")
    ap.add_argument("--roofline-probe", action="store_true",
                    help="only the rotating SpMV launches and calibration copy the counter "
                         "passes profile")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-n", type=int, default=256)
    ap.add_argument("--precision-sigma", type=float, default=1.5)
    ap.add_argument("--precision-max-cycles", type=int, default=4000)
    ap.add_argument("--precision-pcg-iters", type=int, default=5000)
    ap.add_argument("--adjoint-rel-tol", type=float, default=1e-10)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.roofline_probe:
        return roofline_probe(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path "
              "would shard by domain decomposition: one rank per GPU, halo exchange of the "
              "interface x values, scalar all-reductions)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pyopenfoam_amd import PCGSolver
    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.cases import poisson_case
    from pyopenfoam_amd.core.device import current_stream
    from pyopenfoam_amd.differentiable.linear_solver import poisson_solve
    from pyopenfoam_amd.discretisation.fvm import assemble
    from pyopenfoam_amd.mesh.fv_mesh import FvMesh

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = args.n
    t0 = time.perf_counter()
    case = poisson_case(n, jitter=args.jitter, sigma=args.sigma, seed=0,
                        progress=log if n > 128 else None)
    t_case = time.perf_counter() - t0
    N, F = case.n_cells, case.mesh.n_internal_faces
    log(f"case {n}^3: {N} cells, {F} faces, built in {t_case:.1f} s")
    geometry = bench_geometry(case, dev, args.no_cpu)
    mesh = FvMesh.from_case(case, dev)
    pressure = bench_pressure(case, mesh, dev) if not args.no_pressure else None
    t0 = time.perf_counter()
    mat = assemble(case, mesh)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    info = mat.addressing.info
    log(f"handle + assembly {t_asm:.1f} s; slots {info.n_slots}, DIC levels {info.n_dic_levels}, "
        f"DILU levels {info.n_dilu_levels}")

    max_iter = args.max_iter
    if n > 128:  # long solves: report from the solver loop
        mat.addressing.set_progress(lambda it: log(f"  solver loop: {it} iterations enqueued"))
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=max_iter)
    x0 = torch.zeros(N, dtype=torch.float64, device=dev)

    def step():
        mat.invalidate_cache()
        return mat.solve(solver, x0, 1e-300, max_iter)

    for w in range(args.warmup):
        x, iters, res = step()
        torch.cuda.synchronize()
        log(f"warmup {w}: iters {iters}, converged {solver.last_info.converged}")
    torch.cuda.synchronize()
    step_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        x, iters, res = step()
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    converged = bool(solver.last_info.converged)
    r_true = (mat.source - mat.Ax(x)).norm().item() / mat.source.norm().item()
    log(f"solve: {ms:.2f} ms, iters {iters}, converged {converged}, true rel residual {r_true:.3e}")
    if args.history:
        np.save(args.history, np.array(solver.last_info.residual_history))

    # --- SpMV kernel alone: from HBM (rotating copies), resident, ceiling ---
    roof = bench_roofline(case, mat, dev, args)
    spmv_us, spmv_gbs, spmv_bytes = roof["us_hbm"], roof["gbs_hbm"], roof["bytes"]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    if args.copy_probe_mib:
        nb = args.copy_probe_mib * (1 << 20)
        src = torch.empty(nb, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 5 x {args.copy_probe_mib} MiB device copies")

    family = bench_family(mat, dev) if not args.no_family else None

    # --- DIC factor alone ---------------------------------------------------
    from pyopenfoam_amd.solvers import DICPreconditioner
    DICPreconditioner(mat)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        DICPreconditioner(mat)
    e1.record()
    torch.cuda.synchronize()
    dic_ms = e0.elapsed_time(e1) / 5

    # --- adjoint ------------------------------------------------------------
    adjoint = None
    if not args.no_adjoint:
        g = torch.from_numpy(case.gamma).to(dev).requires_grad_(True)
        gb = torch.from_numpy(case.dir_gamma).to(dev).requires_grad_(True)
        w = torch.from_numpy(np.random.default_rng(1).standard_normal(N)).to(dev)
        b = torch.from_numpy(case.b).to(dev)
        dir_data = mesh.dirichlet_data("x0")[1:]
        cells = torch.from_numpy(case.dir_cells.astype(np.int32)).to(dev)
        dd = (cells, torch.from_numpy(case.dir_mag).to(dev), torch.from_numpy(case.dir_delta).to(dev))
        asolver = PCGSolver(tolerance=1e-300, rel_tol=args.adjoint_rel_tol, max_iter=max_iter)
        tf, tb = [], []
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs = poisson_solve(g, gb, b, asolver, mesh, dd)
            J = (w * xs).sum()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            J.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            tf.append(t1 - t0)
            tb.append(t2 - t1)
            g.grad = None
            gb.grad = None
        adjoint = {"forward_ms": round(tf[-1] * 1e3, 2), "backward_ms": round(tb[-1] * 1e3, 2),
                   "ratio": round(tb[-1] / tf[-1], 3), "rel_tol": args.adjoint_rel_tol,
                   "forward_iterations": int(asolver.last_info.iterations)}
        log(f"adjoint: forward {tf[-1] * 1e3:.1f} ms, backward {tb[-1] * 1e3:.1f} ms, "
            f"ratio {tb[-1] / tf[-1]:.3f}")

    # --- mixed precision: FP32 inner PCG in FP64 iterative refinement -------
    mixed = None
    if not args.no_mixed:
        msolver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=max_iter, precision="mixed",
                            inner_rel_tol=args.inner_rel_tol)
        if n <= 128:
            mat.solve(msolver, x0, 1e-300, max_iter)  # warm (packs the FP32 copy)
        else:  # long solves: pack the FP32 copy only
            mat.packed32(False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xm, outer, _ = mat.solve(msolver, x0, 1e-300, max_iter)
        torch.cuda.synchronize()
        mms = (time.perf_counter() - t0) * 1e3
        mi = msolver.last_info
        err = ((xm - x).norm() / x.norm()).item()
        rt = (mat.source - mat.Ax(xm)).norm().item() / mat.source.norm().item()
        mixed = {"ms": round(mms, 2), "speedup_vs_fp64": round(ms / mms, 3),
                 "inner_iterations": mi.inner_iterations, "refinement_steps": outer - 1,
                 "converged": bool(mi.converged), "true_rel_residual": rt,
                 "rel_err_vs_fp64_solution": err, "inner_rel_tol": args.inner_rel_tol}
        log(f"mixed: {mms:.1f} ms ({ms / mms:.2f}x), inner iters {mi.inner_iterations}, "
            f"steps {outer - 1}, true rel residual {rt:.3e}, err vs fp64 {err:.2e}")

    # --- GAMG on the same system ------------------------------------------------
    gamg = None
    if not args.no_gamg:
        from pyopenfoam_amd import GAMGSolver
        gamg = {}
        for name, kw in (("pair_galerkin_jacobi", dict(agglomeration="pair", coarse="galerkin",
                                                        max_levels=20)),
                         ("pair_galerkin_jacobi_mixed", dict(agglomeration="pair", coarse="galerkin",
                                                              max_levels=20, precision="mixed",
                                                              inner_rel_tol=0.1)),
                         ("pair_galerkin_gauss_seidel", dict(agglomeration="pair",
                                                             coarse="galerkin", max_levels=20,
                                                             level_smoother="GaussSeidel")),
                         ("reference", dict()),
                         ("pcg_gamg_precond", dict(pcg=True)),
                         ("pcg_gamg_precond_mixed", dict(pcg=True, precision="mixed"))):
            if args.gamg_variants and name not in args.gamg_variants.split(","):
                continue
            if kw.get("pcg"):  # PCG with one pair/Galerkin V-cycle as the preconditioner
                kwp = {k: v for k, v in kw.items() if k != "pcg"}
                gs = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=args.gamg_max_iter,
                               preconditioner="GAMG", **kwp)
                hier = gs._gamg.hierarchy(mat)
            else:
                gs = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=args.gamg_max_iter, **kw)
                hier = gs.hierarchy(mat)
            mi = args.gamg_max_iter if name != "reference" else 100
            if kw.get("level_smoother"):
                mi = min(mi, 500)  # ~110 ms per cycle: one launch per Gauss-Seidel level
            mat.solve(gs, x0, 1e-300, 2)  # warm (setup of coarse handles and workspace)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xg, cyc, _ = mat.solve(gs, x0, 1e-300, mi)
            torch.cuda.synchronize()
            gms = (time.perf_counter() - t0) * 1e3
            rt = (mat.source - mat.Ax(xg)).norm().item() / mat.source.norm().item()
            extra = {}
            if kw.get("precision") == "mixed" and not kw.get("pcg"):  # cyc counts refinement steps
                extra["refinement_steps"] = int(cyc) - 1
                cyc = gs.last_info.inner_iterations
                ref = gamg.get("pair_galerkin_jacobi")
                if ref is not None:
                    extra["speedup_vs_fp64"] = round(ref["ms"] / gms, 3)
                    extra["rel_err_vs_fp64_solution"] = ((xg - x_fp64).norm() / x_fp64.norm()).item()
            elif not kw.get("pcg"):
                x_fp64 = xg
            if kw.get("pcg") and kw.get("precision") == "mixed":
                ref = gamg.get("pcg_gamg_precond")
                if ref is not None:
                    extra["speedup_vs_fp64"] = round(ref["ms"] / gms, 3)
                    extra["rel_err_vs_fp64_solution"] = ((xg - x_pcg64).norm() / x_pcg64.norm()).item()
            elif kw.get("pcg"):
                x_pcg64 = xg
            if kw.get("level_smoother"):
                extra["gs_level_launches_per_cycle"] = gs_launches_per_cycle(hier, gs)
                extra["us_per_gs_launch"] = round(gms / cyc * 1e3 / extra["gs_level_launches_per_cycle"],
                                                  2)
            gamg[name] = {"ms": round(gms, 2), "cycles": int(cyc), "ms_per_cycle": round(gms / cyc, 3),
                          "converged": bool(gs.last_info.converged), "true_rel_residual": rt,
                          "levels": hier.n_levels, **extra, **{k: v for k, v in kw.items()}}
            log(f"gamg {name}: {gms:.1f} ms, {cyc} cycles ({gms / cyc:.2f} ms/cycle), converged "
                f"{gs.last_info.converged}, true rel residual {rt:.2e}, {hier.n_levels} levels")

    # --- precision: 256^3, lognormal(0, 1.5) -----------------------------------
    precision = None
    if not args.no_precision:
        precision = bench_precision(args, dev)

    # --- batching: many uneven systems in one graph-replayed PCG-DIC ---------
    batch = None
    if not args.no_batch:
        batch = bench_batch(args, dev)

    # --- CPU baseline ---------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        d = mat.diag.cpu().numpy()
        lo = mat.lower.cpu().numpy()
        up = mat.upper.cpu().numpy()
        cpu = cpu_baseline(case, d, lo, up, int(iters), args.cpu_budget)
        log(f"cpu baseline: {cpu['value']} ms per solve ({cpu['sample']})")

    per_iter_bytes = 24 * F + 112 * N
    line = {
        "metric": METRIC,
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "step_ms": [round(v, 3) for v in step_ms],
        "step_ms_median": round(float(np.median(step_ms)), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered hex mesh, lognormal diffusivity, uniform RHS)",
        "config": {"workload": f"PCG-DIC FP64 to rel 1e-8, {n}^3 hex mesh jitter {args.jitter}, "
                               f"permuted+RCM, lognormal(0,{args.sigma}) diffusivity, Dirichlet x0 patch",
                   "n_cells": N, "n_internal_faces": F, "seed": 0, "jitter": args.jitter,
                   "sigma": args.sigma, "max_iter": max_iter,
                   "solver": "PCG", "preconditioner": "DIC", "rel_tol": 1e-8,
                   "tolerance": 1e-300},
        "iterations": int(iters),
        "converged": converged,
        "true_rel_residual": r_true,
        "solve_gbs_algorithmic": round(per_iter_bytes * iters / (ms * 1e-3) / 1e9, 1),
        "dic_factor_ms": round(dic_ms, 3),
        "spmv_us": round(spmv_us, 2),
        "spmv_gbs": round(spmv_gbs, 1),
        "adjoint_over_forward": adjoint["ratio"] if adjoint else None,
        "mixed_precision": mixed,
        "precision_256": precision,
        "adjoint": adjoint,
        "batch": batch,
        "gamg": gamg,
        "roofline": {"bound": "hbm",
                     "kernel": f"{spmv_kernel_name(case)} (pfm_spmv), {n}^3, launches rotating "
                               f"over {ROOFLINE_COPIES} independent copies (HBM-streaming)",
                     "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                     "traffic": roof["traffic"]["bytes_per_launch"] if roof["traffic"] else None,
                     "algorithmic_bytes": spmv_bytes,
                     "us_per_launch": round(spmv_us, 3),
                     "mall_resident": {"us_per_launch": round(roof["us_resident"], 3),
                                       "gbs": round(roof["gbs_resident"], 1)},
                     "read_ceiling_gbs": round(roof["read_ceiling_gbs"], 1),
                     "frac_of_read_ceiling": round(spmv_gbs / roof["read_ceiling_gbs"], 4),
                     "traffic_detail": roof["traffic"]},
        "cpu_baseline": cpu,
        "setup_s": {"case_host": round(t_case, 2), "handle_and_assembly": round(t_asm, 2)},
        "geometry": geometry,
        "pressure_step": pressure,
        "solver_family": family,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

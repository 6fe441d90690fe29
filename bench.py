"""Benchmark: Omniscape circuit solves per second on one MI355X.

Workload (the headline configuration): a 4096 x 4096 float32 resistance map
r = exp(2 z), z a seeded Gaussian random field with a power-law spectrum of
exponent -3 (unit variance), 30 % NoData placed as seeded clustered blobs
(Gaussian-smoothed noise, sigma 32 cells, below its 30 % quantile), sources
1 / r, Omniscape radius 50, block 5, buffer 0, 8 neighbours, block-artifact
correction on. One step is one whole Omniscape run over the map (every target
window factorised, solved and refined to the residual bar, currents summed
into the cumulative map); inputs are resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

prints one JSON line: circuits/s of the mixed-precision GPU path (FP32 MFMA
fronts + FP64 refinement to a relative residual <= 1e-10), with the FP64 GPU
path, the roofline of the dominant kernel and the CPU baseline (the FP64 CPU
restatement of the reference, timed on a bounded sample of the same windows).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

MFMA_F32_PEAK = 157.3e12   # MI355X dense FP32 matrix, FLOP/s
MFMA_F64_PEAK = 78.6e12
HBM_PEAK = 8.0e12          # bytes/s
# HBM traffic per window, from rocprofv3 --pmc FETCH_SIZE (x2, the reading
# of a 256 MiB copy: python bench.py --copy-probe) and WRITE_SIZE (x1) over
# one single-stream run of this workload at 1024 x 1024 (29,147 windows
# factored, 5.05 substitution sweeps each); profiles/ holds the runs.
# One substitution sweep (forward and backward sweeps averaged):
SOLVE_TRAFFIC_PER_WINDOW_SWEEP = 1961950
# The factorisation (assembly and every front level, forward fused):
FACTOR_TRAFFIC_PER_WINDOW = 5294388


def fractal_field(n: int, seed: int, exponent: float = -3.0) -> np.ndarray:
    """Seeded Gaussian random field, power spectrum |k|^exponent, mean 0, variance 1."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (exponent / 2)
    amp[0, 0] = 0.0
    z = np.fft.irfft2(np.fft.rfft2(white) * amp, s=(n, n))
    return (z - z.mean()) / z.std()


def nodata_blobs(n: int, seed: int, sigma: float, fraction: float) -> np.ndarray:
    """Clustered NoData: Gaussian-smoothed white noise below its `fraction` quantile."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    gauss = np.exp(-2.0 * np.pi ** 2 * sigma ** 2 * (kx * kx + ky * ky))
    smooth = np.fft.irfft2(np.fft.rfft2(white) * gauss, s=(n, n))
    return smooth < np.quantile(smooth, fraction)


def workload(n: int = 4096, seed: int = 103, nodata: float = 0.30, sigma: float = 32.0):
    z = fractal_field(n, seed)
    r = np.exp(2.0 * z).astype(np.float32).astype(np.float64)
    r[nodata_blobs(n, seed + 1, sigma, nodata)] = np.nan
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(np.isnan(r), 0.0, 1.0 / r)
    return np.ascontiguousarray(r), np.ascontiguousarray(s)


def pairwise_inputs(n: int, seed: int, npts: int, amp):
    """The pairwise workload: conductance 1 / r on an n x n raster, r = exp(amp z) of the fractal field
    rounded to float32 (amp None: r = exp(N(0, 1)) iid in float64), and npts distinct seeded focal cells."""
    if amp is None:
        r = np.exp(np.random.default_rng(seed).standard_normal((n, n)))
    else:
        r = np.exp(amp * fractal_field(n, seed)).astype(np.float32).astype(np.float64)
    cells = np.random.default_rng(seed + 1000).choice(n * n, npts, replace=False).astype(np.int64)
    return 1.0 / r, cells


def adjoint_inputs(n: int, npairs: int, seed: int):
    """The gradient workload: an n x n lognormal (sigma 1) resistance rounded to float32 and npairs
    seeded (source, ground) cells."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.standard_normal((n, n))).astype(np.float32).astype(np.float64)
    src = rng.choice(n * n, npairs, replace=False)
    dst = rng.choice(n * n, npairs, replace=False)
    return r, src, dst


def large_graph_inputs(n: int = 8192, seed: int = 105, npts: int = 64, nodata: float = 0.10,
                       sigma: float | None = None):
    """The large single-graph workload: r = exp(2 z) of the fractal field in float32, NaN on clustered
    NoData blobs (a `nodata` fraction, sigma n / 128 cells), and npts seeded focal cells among the data
    cells."""
    z = fractal_field(n, seed)
    r = np.exp(2.0 * z).astype(np.float32).astype(np.float64)
    r[nodata_blobs(n, seed + 1, sigma if sigma is not None else n / 128.0, nodata)] = np.nan
    data = np.flatnonzero(~np.isnan(r))
    cells = np.random.default_rng(seed + 1000).choice(data, npts, replace=False).astype(np.int64)
    return np.ascontiguousarray(r), cells


BATCH = 16384  # windows per batch


def log(msg: str) -> None:
    """A progress line on stderr (the JSON result line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def run_gpu(N, d_r, d_s, d_cur, nx, ny, radius, block, precision, tol, timing=True):
    o = N.options(precision, tol, 30, BATCH, timing)
    st = N.SolveStats()
    nt = ctypes.c_int64(0)
    N.check(N.lib().lsc_circuit_omniscape_device(
        ctypes.c_void_p(d_r), ctypes.c_void_p(d_s), ctypes.c_int64(nx), ctypes.c_int64(ny), radius, 0, block, 0, 0,
        1, 0, None, ctypes.byref(o), ctypes.c_void_p(d_cur), None, ctypes.byref(nt), ctypes.byref(st)))
    return st.as_dict()


def host_info() -> dict:
    """The host's CPU model, logical CPUs and RAM (from /proc)."""
    model, ram = platform.processor() or platform.machine(), 0
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
        for line in Path("/proc/meminfo").read_text().splitlines():
            if line.startswith("MemTotal"):
                ram = int(line.split()[1]) * 1024
                break
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count() or 1, "ram_gb": round(ram / 2**30, 1)}


def cpu_baseline(r, s, radius, block, seconds: float):
    """The FP64 CPU oracle (a port of the reference's C++) on full-size windows of the same map: the
    sample sized to about `seconds` of work on all threads, best of three; one thread on a tenth of it."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # marginal rate from two probes (the fixed cost: target list, correction solves)
    t0 = time.perf_counter()
    O.omniscape(r, s, radius, 0, block, False, False, True, False, n_threads=threads, limit=-64)
    t1 = time.perf_counter()
    O.omniscape(r, s, radius, 0, block, False, False, True, False, n_threads=threads, limit=-64 - 32 * threads)
    t2 = time.perf_counter()
    rate = 32 * threads / max((t2 - t1) - (t1 - t0), 1e-3)
    sample = int(min(400000, max(256, rate * seconds / 3)))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = O.omniscape(r, s, radius, 0, block, False, False, True, False, n_threads=threads, limit=-sample)
        times.append(time.perf_counter() - t0)
    dt = min(times)
    one = max(64, sample // (10 * max(threads // 4, 1)))
    t0 = time.perf_counter()
    out1 = O.omniscape(r, s, radius, 0, block, False, False, True, False, n_threads=1, limit=-one)
    dt1 = time.perf_counter() - t0
    return {"value": out["n_targets"] / dt, "unit": "circuits/s", "cores": threads, "kind": "port",
            "sample": f"{out['n_targets']} unclipped radius-{radius} windows of the same map (FP64 multifrontal, "
                      f"{threads} OpenMP threads, best of 3: {', '.join(f'{t:.1f}' for t in times)} s, incl. the "
                      "correction solves)",
            "one_thread": {"value": out1["n_targets"] / dt1, "windows": out1["n_targets"], "s": round(dt1, 1)},
            **host_info()}


def accuracy_crop(N, radius, block, tol, size=448):
    """The cumulative map of a size x size crop of the benchmark map (seeded offset), mixed and FP64
    GPU, against the FP64 CPU oracle: max |difference| over the oracle's max, and the mixed path's
    residuals."""
    from landschaft_amd.circuit import _backend as B
    from oracle import oracle as O

    r, s = workload(4096)
    rng = np.random.default_rng(1103)
    y0, x0 = (int(v) for v in rng.integers(0, 4096 - size, 2))
    r, s = np.ascontiguousarray(r[y0:y0 + size, x0:x0 + size]), np.ascontiguousarray(s[y0:y0 + size, x0:x0 + size])
    ref = O.omniscape(r, s, radius, 0, block, False, False, True, False)["current"]
    top = float(np.max(ref))
    out = {"crop": f"{size}x{size} at ({y0}, {x0})"}
    for prec in ("mixed", "fp64"):
        B.settings.precision, B.settings.tol = prec, tol
        cur = B.circuit_omniscape(r, s, radius, 0, block, False, False, True, False)["current"]
        st = B.last_stats
        out[prec] = {"cum_max_rel_err_vs_cpu_fp64": float(np.max(np.abs(cur - ref)) / top),
                     "resid_max": st["resid_max"], "resid_median": st["resid_median"], "windows": st["n_solved"]}
    B.settings.precision = "mixed"
    return out


def _phase(st):
    return {"ms_analysis": round(st["ms_fill"], 2), "ms_factor": round(st["ms_factor"], 2),
            "ms_solve": round(st["ms_solve"], 2), "resid_max": st["resid_max"], "resid_median": st["resid_median"],
            "refine_max": st["refine_max"], "refine_mean": round(st["refine_mean"], 2), "n_solves": st["n_solved"],
            "n_fallback": st["n_fallback"]}


def pairwise_leg(B, n, four, npts, seed, amp, maps_too, cpu=False):
    """Circuitscape pairwise on one n x n graph: mixed and FP64 GPU (and optionally the CPU oracle)."""
    g, cells = pairwise_inputs(n, seed, npts, amp)
    ex = np.zeros((npts, npts), np.uint8)
    out = {"workload": f"pairwise {n}x{n} {'4' if four else '8'}-neighbour, {npts} focal cells "
                       f"({npts * (npts - 1) // 2} pairs), seed {seed}"}
    res = {}
    for prec in ("mixed", "fp64"):
        B.settings.precision = prec
        t0 = time.perf_counter()
        R = B.circuitscape_pairwise(g, None, cells, ex, False, four, {})["resistances"]
        out[prec] = {"s_wall": round(time.perf_counter() - t0, 3), **_phase(B.last_stats)}
        res[prec] = R
    off = ~np.eye(npts, dtype=bool)
    out["max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(res["mixed"] - res["fp64"])[off] / res["fp64"][off]))
    if maps_too:
        B.settings.precision = "mixed"
        t0 = time.perf_counter()
        Ra = B.circuitscape_pairwise(g, None, cells, ex, False, four, {"cum": True})["resistances"]
        out["mixed_all_pairs_cum_map"] = {"s_wall": round(time.perf_counter() - t0, 3), **_phase(B.last_stats),
                                          "max_rel_diff_vs_shortcut": float(
                                              np.max(np.abs(Ra - res["mixed"])[off] / res["mixed"][off]))}
    if cpu:
        from oracle import oracle as O

        t0 = time.perf_counter()
        Rc = O.cs_pairwise(g, None, cells, ex, False, four, {"cum": True, "volt_maps": True})["resistances"]
        out["cpu_oracle_fp64"] = {"s_wall": round(time.perf_counter() - t0, 3),
                                  "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                                  "max_rel_diff_gpu_mixed": float(
                                      np.max(np.abs(res["mixed"] - Rc)[off] / Rc[off]))}
    B.settings.precision = "mixed"
    return out


def network_leg(B, n, npts, seed, cpu=False):
    """Circuitscape's network pairwise on the edge list of an n x n 4-neighbour grid (conductance
    (g_a + g_b) / 2 per edge: the Laplacian of the raster with four_neighbors), every pair solved; the
    same pairs on the raster path (geometric ordering, shortcut) as a cross-check, and optionally the
    CPU oracle's raster pairwise."""
    r = np.exp(2.0 * fractal_field(n, seed)).astype(np.float32).astype(np.float64)
    g = 1.0 / r
    idx = np.arange(n * n, dtype=np.int64).reshape(n, n)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    w = (g.ravel()[a] + g.ravel()[b]) / 2
    cells = np.random.default_rng(seed + 1000).choice(n * n, npts, replace=False).astype(np.int64)
    ex = np.zeros((npts, npts), np.uint8)
    out = {"workload": f"network pairwise: {n}x{n} 4-neighbour grid as an edge list ({n * n} nodes, {a.size} "
                       f"edges), {npts} focal nodes ({npts * (npts - 1) // 2} pairs, all solved), seed {seed}"}
    B.settings.precision = "mixed"
    t0 = time.perf_counter()
    Rn = B.network_pairwise(a.astype(np.int32), b.astype(np.int32), w, n * n, cells.astype(np.int32), ex, {})
    out["mixed"] = {"s_wall": round(time.perf_counter() - t0, 3), **_phase(B.last_stats)}
    t0 = time.perf_counter()
    Rr = B.circuitscape_pairwise(g, None, cells, ex, False, True, {})["resistances"]
    out["raster_path_mixed"] = {"s_wall": round(time.perf_counter() - t0, 3), **_phase(B.last_stats)}
    off = ~np.eye(npts, dtype=bool)
    out["max_rel_diff_network_vs_raster"] = float(np.max(np.abs(Rn["resistances"] - Rr)[off] / Rr[off]))
    if cpu:
        from oracle import oracle as O

        t0 = time.perf_counter()
        Rc = O.cs_pairwise(g, None, cells, ex, False, True, {})["resistances"]
        out["cpu_oracle_fp64_raster"] = {"s_wall": round(time.perf_counter() - t0, 3),
                                         "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                                         "max_rel_diff_gpu_network": float(
                                             np.max(np.abs(Rn["resistances"] - Rc)[off] / Rc[off]))}
    return out


def adjoint_leg(B, n, npairs, seed):
    """J = sum of effective resistances and dJ/dr on an n x n 8-neighbour lognormal surface."""
    r, src, dst = adjoint_inputs(n, npairs, seed)
    out = {"workload": f"adjoint {n}x{n} 8-neighbour lognormal sigma 1, {npairs} pairs, seed {seed}"}
    grads = {}
    for prec in ("mixed", "fp64"):
        B.settings.precision = prec
        t0 = time.perf_counter()
        J, R, grad = B.effective_resistance_grad(r, src, dst)
        out[prec] = {"s_wall": round(time.perf_counter() - t0, 3), "J": J, **_phase(B.last_stats)}
        grads[prec] = grad
    out["grad_max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(grads["mixed"] - grads["fp64"]))
                                                   / np.abs(grads["fp64"]).max())
    B.settings.precision = "mixed"
    return out


def large_graph_leg(B, npairs=16):
    """The 8192 x 8192 single graph (60 M unknowns, 10 % NoData, 4 neighbours): all-to-one over 64
    focal cells with the cumulative map, mixed and FP64, and dJ/dr of J = sum of the effective
    resistances of npairs pairs of focal cells (one calibration step)."""
    from landschaft_amd.circuit import circuitscape as cs

    t0 = time.perf_counter()
    r, cells = large_graph_inputs()
    gen = time.perf_counter() - t0
    pm = np.zeros(r.shape)
    pm.flat[cells] = np.arange(1, cells.size + 1)
    hab = np.where(np.isnan(r), cs.NODATA, r)
    out = {"workload": "all-to-one 8192x8192 fractal exp(2z) f32, 10% nodata blobs, 4-neighbour, 64 focal cells, "
                       f"cum map; dJ/dr of {npairs} focal pairs", "gen_s": round(gen, 1)}
    cums = {}
    for prec in ("mixed", "fp64"):
        B.settings.precision = prec
        t0 = time.perf_counter()
        res = cs.circuitscape_one_to_all(hab, pm, all_to_one=True, four_neighbors=True, cum_map=True)
        out[prec] = {"s_wall": round(time.perf_counter() - t0, 2), **_phase(B.last_stats)}
        cums[prec] = res.cum_current
    top = float(np.max(cums["fp64"]))
    out["cum_max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(cums["mixed"] - cums["fp64"])) / top)
    del cums
    B.settings.precision = "mixed"
    t0 = time.perf_counter()
    from scipy import ndimage

    lab, _ = ndimage.label(~np.isnan(r))  # 4-connected components of the data cells
    same = cells[lab.flat[cells] == lab.flat[cells[0]]]
    src, dst = same[:npairs], same[npairs:2 * npairs]
    J, _, grad = B.effective_resistance_grad(r, src, dst, True)
    out["grad_mixed"] = {"s_wall": round(time.perf_counter() - t0, 2), "J": J, "pairs": int(src.size),
                         **_phase(B.last_stats),
                         "grad_min": float(np.min(grad)), "cells": int(grad.size)}
    return out


def tune(a) -> int:
    """A small autotuner: every combination of the window engine's tunables on one map, best first. The
    library reads LSC_ND_LEAF and LSC_BATCH_STREAMS at each call; the batch size is a solve option."""
    import torch

    from landschaft_amd import _native as N

    global BATCH
    n = a.size if a.size != 4096 else 2048
    r, s = workload(n)
    dev = torch.device("cuda:0")
    tr, ts = torch.from_numpy(r).to(dev), torch.from_numpy(s).to(dev)
    tc = torch.empty((n, n), dtype=torch.float64, device=dev)
    args = (N, tr.data_ptr(), ts.data_ptr(), tc.data_ptr(), n, n, a.radius, a.block)
    results = []
    saved = {k: os.environ.get(k) for k in ("LSC_ND_LEAF", "LSC_BATCH_STREAMS")}
    try:
        for leaf in (24, 32, 40):
            for streams in (1, 2, 3, 4):
                for batch in (8192, 16384):
                    os.environ["LSC_ND_LEAF"], os.environ["LSC_BATCH_STREAMS"] = str(leaf), str(streams)
                    BATCH = batch
                    run_gpu(*args, "mixed", a.tol, timing=False)  # warm-up (plans, allocations)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    st = run_gpu(*args, "mixed", a.tol, timing=False)
                    torch.cuda.synchronize()
                    dt = time.perf_counter() - t0
                    results.append({"leaf": leaf, "streams": streams, "batch": batch,
                                    "circuits_per_s": round(st["n_targets"] / dt, 1), "resid_max": st["resid_max"]})
                    print(json.dumps(results[-1]), file=sys.stderr, flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    results.sort(key=lambda d: -d["circuits_per_s"])
    print(json.dumps({"tune": f"omniscape {n}x{n} radius {a.radius} block {a.block}", "best": results[0],
                      "all": results}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--radius", type=int, default=50)
    ap.add_argument("--block", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the pairwise and adjoint graph workloads")
    ap.add_argument("--no-large", action="store_true", help="skip the 8192^2 single-graph workload")
    ap.add_argument("--batch", type=int, default=16384, help="Omniscape windows per batch")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 256 MiB buffer three times (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--tune", action="store_true",
                    help="time the window engine's tunables (nested-dissection leaf size, batches in flight, batch "
                         "size) on the --size map (2048 unless given) and print the fastest setting")
    a = ap.parse_args()
    if a.tune:
        return tune(a)
    if a.copy_probe:
        import torch

        buf = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        for _ in range(3):
            buf.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": buf.numel() * 4, "copies": 3}))
        return 0
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the path shards by row bands "
                                   "of target windows, see DESIGN.md)"}))
        return 2
    import torch

    from landschaft_amd import _native as N

    global BATCH
    BATCH = a.batch

    n = a.size
    t0 = time.perf_counter()
    r, s = workload(n)
    gen_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    tr = torch.from_numpy(r).to(dev)
    ts = torch.from_numpy(s).to(dev)
    tc = torch.empty((n, n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    args = (N, tr.data_ptr(), ts.data_ptr(), tc.data_ptr(), n, n, a.radius, a.block)
    log(f"map {n}x{n} generated in {gen_s:.1f} s; {a.warmup} warm-up and {a.steps} timed steps")
    for _ in range(a.warmup):
        run_gpu(*args, "mixed", a.tol)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        stats.append(run_gpu(*args, "mixed", a.tol))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    st = stats[-1]
    value = st["n_targets"] / dt
    cum_mixed = tc.cpu().numpy()

    # Algorithmic per-window figures of the reference's ordering (16-cell
    # nested-dissection leaves): front flops sum np^3/3 + np^2 nu + np nu^2,
    # and L as m x np column blocks in FP32 read once per substitution sweep.
    plan = N.window_plan_stats(a.radius, 0, False, True)

    # The factor's algorithmic bytes: L written once (FP32, the reference's
    # ordering) and the window's conductances read once (W^2 FP64). Its
    # intensity, factor_flops / those bytes, is about 11 FLOP/B, below the
    # FP32 matrix ridge of 157.3 / 8 = 19.7 FLOP/B: the HBM roof bounds it.
    w_side = 2 * a.radius + 1
    factor_bytes = plan["l_unpacked"] * 4 + w_side * w_side * 8

    def roofline_of(st):
        kernels = {}
        f_flops = plan["factor_flops"] * st["n_factored"]
        f_s = st["ms_factor"] / 1e3
        kernels["k_factor"] = {"bound": "hbm", "achieved": st["n_factored"] * factor_bytes / f_s / 1e9,
                               "peak": HBM_PEAK / 1e9, "unit": "GB/s", "launches": st["n_batches"],
                               "ms_total": st["ms_factor"], "bytes_per_circuit": factor_bytes,
                               "flops_per_circuit": plan["factor_flops"],
                               "flop_per_byte": round(plan["factor_flops"] / factor_bytes, 2),
                               "mfma_tflops": round(f_flops / f_s / 1e12, 3),
                               "mfma_frac": round(f_flops / f_s / MFMA_F32_PEAK, 4)}
        # sweeps: the back substitution after the fused factor-forward, then a
        # forward + back sweep per refinement step
        sweeps = st["n_factored"] * (1 + 2 * st["refine_mean"])
        l_bytes = plan["l_unpacked"] * 4
        s_s = st["ms_solve"] / 1e3
        kernels["k_solve"] = {"bound": "hbm", "achieved": sweeps * l_bytes / s_s / 1e9, "peak": HBM_PEAK / 1e9,
                              "unit": "GB/s", "launches": st["n_solve_launches"], "ms_total": st["ms_solve"],
                              "bytes_per_sweep": l_bytes}
        dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
        kd = kernels[dom]
        # one launch of k_solve = one substitution sweep over a batch of windows
        per_batch = st["n_factored"] / max(st["n_batches"], 1)
        per_launch = {"k_factor": FACTOR_TRAFFIC_PER_WINDOW, "k_solve": SOLVE_TRAFFIC_PER_WINDOW_SWEEP}
        traffic = round(per_launch[dom] * per_batch)
        roof = {"kernel": dom, "bound": kd["bound"], "achieved": round(kd["achieved"], 3), "peak": kd["peak"],
                "unit": kd["unit"], "frac": round(kd["achieved"] / kd["peak"], 4), "traffic": traffic,
                "concurrent_batches": st["n_streams"],
                "traffic_note": ("HBM bytes per launch: " + ("the factorisation of a batch of windows"
                                 if dom == "k_factor" else "one substitution sweep over a batch of windows")
                                 + " (rocprofv3 PMC, copy-calibrated)")}
        if dom == "k_factor":
            roof["mfma_tflops"], roof["mfma_frac"] = kd["mfma_tflops"], kd["mfma_frac"]
        return roof, kernels

    # the timed region: launch durations are timed on each batch's stream
    # while the other batches run beside it (per-launch rates, overlapping)
    roofline, kernels = roofline_of(st)
    if st["n_streams"] > 1:
        roofline["achieved_note"] = (f"per launch over the timed region, {st['n_streams']} batches in flight on "
                                     "their own streams, so launches overlap; single_stream: the same run with one "
                                     "batch in flight (no overlap), untimed")
        os.environ["LSC_BATCH_STREAMS"] = "1"
        try:
            st1 = run_gpu(*args, "mixed", a.tol)
        finally:
            del os.environ["LSC_BATCH_STREAMS"]
        r1, k1 = roofline_of(st1)
        roofline["single_stream"] = {k: r1[k] for k in ("kernel", "bound", "achieved", "frac", "traffic")}
        roofline["single_stream"]["phase_ms"] = {k[3:]: round(v, 2) for k, v in st1.items() if k.startswith("ms_")}

    log(f"mixed: {value:.0f} circuits/s")
    fp64 = None
    if not a.no_fp64:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st64 = run_gpu(*args, "fp64", a.tol)
        torch.cuda.synchronize()
        d64 = time.perf_counter() - t0
        cum64 = tc.cpu().numpy()
        top = np.nanmax(np.abs(cum64))
        fp64 = {"value": st64["n_targets"] / d64, "ms": d64 * 1e3, "resid_max": st64["resid_max"],
                "cum_max_rel_diff_vs_mixed": float(np.nanmax(np.abs(cum_mixed - cum64)) / top),
                "refine_mean": st64["refine_mean"],
                "phase_ms": {k[3:]: round(v, 2) for k, v in st64.items() if k.startswith("ms_")}}
    log("CPU baseline")
    cpu = None if a.no_cpu else cpu_baseline(r, s, a.radius, a.block, a.cpu_seconds)
    graph = None
    if not a.no_graph:
        from landschaft_amd.circuit import _backend as B

        B.settings.tol = a.tol
        graph = {}
        log("graph workloads: pairwise 256^2")
        graph["pairwise_256"] = pairwise_leg(B, 256, False, 4, 101, None, True, cpu=not a.no_cpu)
        log("pairwise 2048^2")
        graph["pairwise_2048"] = pairwise_leg(B, 2048, True, 16, 102, 2.0, True)
        log("adjoint 1024^2")
        graph["adjoint_1024"] = adjoint_leg(B, 1024, 32, 104)
        log("network 1024^2")
        graph["network_1024"] = network_leg(B, 1024, 8, 105, cpu=not a.no_cpu)
        if not a.no_large:
            log("large graph 8192^2")
            graph["large_8192"] = large_graph_leg(B)
    log("accuracy crop against the CPU oracle")
    accuracy = None if a.no_cpu else accuracy_crop(N, a.radius, a.block, a.tol)
    line = {
        "metric": "circuit solves/sec at 4096^2 @ rel resid <=1e-10 (Omniscape windows, mixed-precision GPU)",
        "value": round(value, 2), "unit": "circuits/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-mfma fronts + f64 refinement", "data": "synthetic",
        "config": {"workload": f"omniscape {n}x{n} fractal exp(2z) f32, 30% nodata blobs, radius {a.radius}, "
                               f"block {a.block}, 8-neighbour, correction on",
                   "radius": a.radius, "block_size": a.block, "tol": a.tol, "seed": 103},
        "n_targets": st["n_targets"], "resid_max": st["resid_max"], "resid_median": st["resid_median"],
        "n_over_1e10": st["n_over_1e10"], "n_fallback": st["n_fallback"], "accuracy_vs_cpu_fp64": accuracy,
        "refine_mean": round(st["refine_mean"], 3), "refine_max": st["refine_max"],
        "phase_ms": {k[3:]: round(v, 2) for k, v in st.items() if k.startswith("ms_")},
        "kernels": kernels, "roofline": roofline, "fp64_gpu": fp64, "cpu_baseline": cpu, "graph": graph,
        "gen_seconds": round(gen_s, 1),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

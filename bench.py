#!/usr/bin/env python3
"""Scatter events/s of the DS electron transport on one MI355X.

Workload: plain DS mode, semi-infinite Cu (Z=29), E0 = 20 keV tracked to
Emin = 1 keV, s_tilt = 70 deg, Wc = 50 eV, tol_E = tol_W = 1e-3, 2^22 electrons,
Philox seed 1.  One step = one mcd_forward launch over the whole electron batch
(tables built beforehand, outputs resident in HBM), at the launch shape the
autotuner picks before the timed region (guard placement, workgroup size,
waves per SIMD; records are identical for every shape).  The channel-pick
table and the W tables are built once outside the timed region.

It also reports the 64-configuration batched launch against 64 sequential
launches (same GPU, same Philox keys).  The second half of the metric, the
backward/forward time ratio, is measured
on the differentiation workload: Cu 20 keV -> 1 keV, s_tilt 70 deg, DS
(Wc 50 eV) and continuous JL, 2^24 electrons, seed 4; forward = one
mcd_forward launch at its autotuned shape, backward = one mcd_backward launch
(vector-Jacobian product for the yield) over the same electrons.

Prints one JSON line.  With --gpus N > 1 it exits: the benchmark measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mc-discrete_amd"))

METRIC = "scatter events/s on 1 MI355X (DS, Cu 20 keV, 70°); backward/forward time ratio"
# algorithmic FP32 operations per DS scatter event for Cu (add/sub/mul/div/cmp = 1,
# FMA = 2, sqrt/rsqrt/log/sin = 1; Philox integer work excluded)
F_OPS_CU_DS = 243
PEAK_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--electrons", type=int, default=1 << 22)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target wall time of each CPU-baseline sample (samples hold at least 2^14 electrons)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=0)
    ap.add_argument("--guard", choices=("inline", "deferred", "service"), default=None,
                    help="fix the FP64 guard placement (with --block/--waves-per-simd: skips the autotuner)")
    ap.add_argument("--no-backward", action="store_true", help="skip the backward/forward ratio")
    ap.add_argument("--bwd-electrons", type=int, default=1 << 24)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-configuration batched comparison")
    ap.add_argument("--no-two-stream", action="store_true",
                    help="skip the two-stream pipelined throughput (reported beside value)")
    ap.add_argument("--config-sweep", action="store_true",
                    help="instead of the headline line, print one JSON line per parity configuration "
                         "(FP64 Si 10 keV, DS Cu 20 keV, continuous Au 30 keV x 3 Bethe laws x 2 tilts)")
    ap.add_argument("--valu-ceiling", action="store_true",
                    help="instead of the headline line, time the issue-rate calibration kernels (v_fma_f32, "
                         "v_mad_u64_u32, v_bitop3_b32 at full occupancy) and print one JSON line")
    ap.add_argument("--precision", action="store_true",
                    help="instead of the headline line, run the precision study (FP32 vs the all-FP64 GPU run of "
                         "the same streams; DS Cu 20 keV and continuous Au 30 keV x 3 Bethe laws x 2 tilts) and "
                         "print one JSON line per workload")
    ap.add_argument("--cli", action="store_true",
                    help="instead of the headline line, time the doScatter.py command line end to end on the "
                         "headline configuration (input file, keyed tables, chunked launches, HDF5 store): "
                         "electron records of 2^22 electrons, then records and four scatter fields of 2^12 "
                         "electrons, each with the tables built and then read back; one JSON line")
    ap.add_argument("--fixed-shapes", action="store_true",
                    help="skip every autotuner: the headline runs served 1024/8 (unless --block/--guard say "
                         "otherwise), the differentiation forward and the continuous runs the library default, "
                         "the batch 128 threads, 128 pick nodes, guard inline (for kernel profiles)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed steps, run one device-to-device copy of this size (the reference "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE readings)")
    return ap.parse_args()


def cpu_cores():
    """(threads used, host summary): every core this process may run on
    (sched_getaffinity), or fewer when OMP_NUM_THREADS states the process's
    CPU share (a shared GPU host gives each GPU's jobs such a share); the
    summary reports os.cpu_count() and the affinity beside it."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    n = aff
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n), {"os_cpu_count": os.cpu_count(), "affinity": aff, "omp_num_threads": env}


def cpu_baselines(tr, seconds, seed):
    """The FP64 C oracle (OpenMP) and the Python reference-path restatement on
    bounded samples of the same workload, timed on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (the checker: CPU-baseline leg only)
    import baseline_b1 as B1  # noqa: E402

    gm, gg = tr.grids
    tm = O.Table(gm.Es, gm.Ws, tr.cdf(0))
    tg = O.Table(gg.Es, gg.Ws, tr.cdf(1))
    threads, host = cpu_cores()
    # C oracle: at least 2^14 electrons, grown until the sample takes about `seconds`
    n, ev, dt = 1 << 14, 0, 0.0
    while True:
        t = time.perf_counter()
        _, _, ev = O.run(tr.cfg, tm, tg, seed, 0, n, nthreads=threads)
        dt = time.perf_counter() - t
        if dt >= 0.8 * seconds or n >= (1 << 24):
            break
        n = int(n * min(8.0, max(1.25, 1.05 * seconds / max(dt, 1e-3))))
    c_base = {"value": ev / dt, "unit": "events/s", "cores": threads, "kind": "port", "host": host,
              "sample": f"FP64 C oracle (OpenMP), electrons 0..{n - 1} of the same workload, {ev} events in {dt:.1f} s"}
    # Python reference-path restatement under multiprocessing (4 trajectories per
    # job): at least 2^14 electrons, sized from a short serial probe to take about `seconds`
    P = B1.params_from(tr.cfg, tm, tg, seed=seed)
    t = time.perf_counter()
    B1.run_serial(P, 0, 16)
    per_el = (time.perf_counter() - t) / 16
    n1 = max(1 << 14, int(seconds * threads / max(per_el, 1e-6)))
    ev1, dt1, w = B1.run_pool(P, n1, workers=threads)
    b1 = {"value": ev1 / dt1, "unit": "events/s", "cores": w, "kind": "port", "host": host,
          "sample": f"Python/NumPy reference-path restatement, Pool({w}) x 4 trajectories per job, "
                    f"electrons 0..{n1 - 1}, {ev1} events in {dt1:.1f} s"}
    # the same restatement on one worker (the reference's num_workers = 1) and
    # as the reference ships it (every job carries the tables, one per shell)
    ns = max(16, int(0.5 * seconds / max(per_el, 1e-6)))
    t = time.perf_counter()
    evs, _ = B1.run_serial(P, 0, ns)
    dts = time.perf_counter() - t
    nsh = max(4 * threads, n1 // 4)
    evh, dth, wh = B1.run_pool_shipped(P, nsh, workers=threads, n_gryz_tables=int(tr.cfg.mat.n_shells))
    b1["one_worker"] = {"value": evs / dts, "cores": 1,
                        "sample": f"electrons 0..{ns - 1}, {evs} events in {dts:.1f} s"}
    b1["as_shipped"] = {"value": evh / dth, "cores": wh,
                        "sample": f"Pool({wh}), one apply_async per 4 trajectories carrying the Moller table and "
                                  f"{int(tr.cfg.mat.n_shells)} Gryzinski tables, electrons 0..{nsh - 1}, "
                                  f"{evh} events in {dth:.1f} s"}
    return c_base, b1


def backward_ratio(n, reps=3, tune=True):
    """Forward and backward kernel times (HIP events on the launch stream) on
    the differentiation workload, DS and continuous JL; the forward runs at
    the launch shape the autotuner picks for that workload."""
    import torch

    from mcd import _abi
    from mcd.autograd import BackscatterSpec

    dev = torch.device("cuda", 0)
    out = {"workload": f"Cu bulk 20 keV -> 1 keV, s_tilt 70 deg, {n} electrons, seed 4; "
                       "backward = vector-Jacobian product of the yield w.r.t. (Wc, s_tilt, J, density)"}
    els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    g = torch.zeros(65, dtype=torch.float32, device=dev)
    g[0] = 1.0
    acc = torch.zeros(1 + _abi.NGRAD, dtype=torch.float64, device=dev)
    for key, mode in (("DS", "DS"), ("cont_JL", "cont")):
        spec = BackscatterSpec(mode=mode, n=n, seed=4)
        tr = spec.transport(spec.params().tolist())
        shape = tr.tune(n, seed=4, reps=1) if tune else dict(block=0, waves_per_simd=0, guard=None, refill_period=1)
        tr.launch(n, els, cnt, seed=4)
        tr.vjp(n, g, seed=4, out=acc)
        torch.cuda.synchronize(dev)
        tf, tb = [], []
        for _ in range(reps):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            e[0].record()
            tr.launch(n, els, cnt, seed=4)
            e[1].record()
            e[2].record()
            tr.vjp(n, g, seed=4, out=acc)
            e[3].record()
            torch.cuda.synchronize(dev)
            tf.append(e[0].elapsed_time(e[1]))
            tb.append(e[2].elapsed_time(e[3]))
        f, b = min(tf), min(tb)
        out[key] = {"forward_ms": f, "backward_ms": b, "ratio": b / f,
                    "forward_shape": {k: shape[k] for k in ("block", "waves_per_simd", "guard", "refill_period")}}
    out["ratio_max"] = max(out["DS"]["ratio"], out["cont_JL"]["ratio"])
    return out


def two_stream(tr, n, steps, seed):
    """The serial loop's steps with consecutive launches alternating between
    two HIP streams (default priority), each with its own record and counter
    buffers: the next batch starts on the CUs that the previous launch's last
    trajectories leave idle.  Same launches, same work per step."""
    import torch

    from mcd import _abi
    dev = torch.device("cuda", 0)
    ss = [torch.cuda.Stream(device=dev) for _ in range(2)]
    bufs = [(torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev),
             torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)) for _ in range(2)]

    def go(k):
        els, cnt = bufs[k % 2]
        with torch.cuda.stream(ss[k % 2]):
            tr.launch(n, els, cnt, seed=seed, stream=ss[k % 2])

    go(0)
    go(1)
    torch.cuda.synchronize(dev)
    for _, cnt in bufs:
        cnt.zero_()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for k in range(steps):
        go(k)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    events = sum(int(cnt[_abi.CNT_EVENTS].item()) for _, cnt in bufs)
    return {"events_per_s": events / wall, "ms_per_step": wall / steps * 1e3, "streams": 2,
            "note": "consecutive batches overlap: reported beside value, not as value"}


def batched_ratio(reps=3, tune=True):
    """64 configurations: 64 sequential mcd_forward launches vs one autotuned,
    graph-replayed batched launch (same Philox keys)."""
    import torch

    from mcd import _abi, autotune, engine, workloads

    dev = torch.device("cuda", 0)
    cfgs = workloads.batched_configs()
    # the 64 configurations' tables: one library build per configuration
    # (mcd_tables_build) against one batched build (mcd_tables_build_batch)
    t = time.perf_counter()
    one_by_one = [engine.Transport(c["material"], c["E0"], c["Emin"], Wc=c["Wc"], s_tilt=c["s_tilt"], mode=c["mode"],
                                   bethe_model=c["bethe_model"], tol_E=c["tol"], tol_W=c["tol"],
                                   grid_search="library") for c in cfgs]
    torch.cuda.synchronize(dev)
    t_one = time.perf_counter() - t
    del one_by_one
    t = time.perf_counter()
    trs = workloads.build_transports(cfgs, batch_tables=True)
    torch.cuda.synchronize(dev)
    t_batch_tables = time.perf_counter() - t
    # the same batched build with the grid searches on host threads
    os.environ["MCD_GRID_SEARCH"] = "host"
    try:
        t = time.perf_counter()
        trs_host = workloads.build_transports(cfgs, batch_tables=True)
        torch.cuda.synchronize(dev)
        t_batch_host = time.perf_counter() - t
    finally:
        os.environ.pop("MCD_GRID_SEARCH", None)
    del trs_host
    # the grid searches alone: every DS table in one device launch
    # (mcd_grid_search_batch) against the host search of each on one core
    from mcd.material import Material
    from mcd.tables import gryz_grid, moller_grid
    grids = []
    for c in cfgs:
        if c["mode"] == "DS":
            m = Material(c["material"])
            grids += [moller_grid(m, c["E0"], c["Emin"], c["Wc"]), gryz_grid(m, c["E0"], c["Emin"])]
    tol = cfgs[0]["tol"]
    t = time.perf_counter()
    engine.grid_search_batch(grids, tol, tol)
    t_search_dev = time.perf_counter() - t
    t = time.perf_counter()
    for g in grids:
        g.build_grids_library(tol, tol)
    t_search_host = time.perf_counter() - t
    items = [(tr, c["n"]) for tr, c in zip(trs, cfgs)]
    els = [torch.empty(c["n"] * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev) for c in cfgs]
    cnt = torch.zeros(len(cfgs) * _abi.NCOUNTERS, dtype=torch.int64, device=dev)

    def seq():
        for i, (tr, c) in enumerate(zip(trs, cfgs)):
            tr.launch(c["n"], els[i], cnt[i * _abi.NCOUNTERS:(i + 1) * _abi.NCOUNTERS], seed=3, config_id=i)

    def best(f):
        f()
        torch.cuda.synchronize(dev)
        t_best = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize(dev)
            t = time.perf_counter()
            f()
            torch.cuda.synchronize(dev)
            t_best = min(t_best, time.perf_counter() - t)
        return t_best * 1e3

    t_seq = best(seq)
    tuned = (autotune.tune_batch(items, seed=3, reps=1, use_cache=False) if tune else
             dict(block=128, pick_nodes=128, guard="inline", refill_period=1, ms={}))
    b = engine.Batch(items, pick_nodes=tuned["pick_nodes"])
    t_bat = best(lambda: b.launch(seed=3, use_graph=True, block=tuned["block"], guard=tuned["guard"],
                                  refill_period=tuned.get("refill_period", 1)))
    return {"workload": "64 configurations {Al,Si,Cu,Au} x E0 {5,10,20,30} keV x tilt {0,70} deg x "
                        "{DS, continuous JL}, Emin 0.5 keV, Wc 100 eV, tol 1e-3, "
                        f"{sum(c['n'] for c in cfgs)} electrons (log-uniform 2^14..2^20, seed 3)",
            "sequential_ms": t_seq, "batched_ms": t_bat, "ratio": t_seq / t_bat,
            "tables": {"per_configuration_ms": t_one * 1e3, "batched_ms": t_batch_tables * 1e3,
                       "batched_host_search_ms": t_batch_host * 1e3,
                       "grid_search": {"tables": len(grids), "device_one_launch_ms": t_search_dev * 1e3,
                                       "host_one_core_ms": t_search_host * 1e3},
                       "note": "DS tables of the 64 configurations from tol_E = tol_W = 1e-3: grid search, FP64 "
                               "CDF rows, row records and channel-pick tables; mcd_tables_build per configuration "
                               "vs one mcd_tables_build_batch (grid searches in one device launch; with "
                               "MCD_GRID_SEARCH=host on host threads)"},
            "tuned": {"block": tuned["block"], "pick_nodes": tuned["pick_nodes"], "guard": tuned["guard"],
                      "refill_period": tuned.get("refill_period", 1), "candidates_ms": tuned["ms"]}}


def config_sweep(steps, tune=True):
    """Throughput of the configurations the parity tests run: the all-FP64
    kernel on Si 10 keV (2,000 electrons, seed 0), the DS headline (Cu 20 keV,
    70 deg, 2^22 electrons, seed 1) and continuous Au 30 keV -> 1 keV with the
    classical, JL and explicit Bethe laws at 0 and 70 deg (2^22 electrons,
    seed 2), and the DS headline with each of the other elastic models, each
    at its autotuned launch shape (FP64: library default)."""
    import torch

    from mcd import _abi, engine
    dev = torch.device("cuda", 0)
    runs = [("FP64 DS Si 10 keV -> 1 keV, 70 deg, Wc 100 eV, tol 1e-2", 2000, 0,
             dict(material="Si", E0=10000.0, Emin=1000.0, Wc=100.0, s_tilt=70.0, tol_E=1e-2, tol_W=1e-2), True),
            ("DS Cu 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3", 1 << 22, 1,
             dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, s_tilt=70.0), False)]
    for law in ("classical", "JL", "explicit"):
        for tilt in (0.0, 70.0):
            runs.append((f"continuous {law} Au 30 keV -> 1 keV, {tilt:g} deg", 1 << 22, 2,
                         dict(material="Au", E0=30000.0, Emin=1000.0, s_tilt=tilt, mode="cont", bethe_model=law),
                         False))
    # the other elastic models (crossSections.py:63-147) on the DS headline
    for model in ("Ruth_vanilla_wDefl", "Ruth_nigram", "Ruth_nigram_wDefl"):
        runs.append((f"DS Cu 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3, elastic {model}", 1 << 22, 1,
                     dict(material="Cu", E0=20000.0, Emin=1000.0, Wc=50.0, s_tilt=70.0, elastic=model), False))
    for name, n, seed, kw, fp64 in runs:
        tr = engine.Transport(**kw)
        prec = _abi.PREC_FP64 if fp64 else _abi.PREC_FP32
        if not fp64 and tune:
            tr.tune(n, seed=seed, reps=1, use_cache=False)
        els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
        tr.launch(n, els, cnt, seed=seed, precision=prec)
        torch.cuda.synchronize(dev)
        cnt.zero_()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
        for k in range(steps):
            ev[2 * k].record()
            tr.launch(n, els, cnt, seed=seed, precision=prec)
            ev[2 * k + 1].record()
        torch.cuda.synchronize(dev)
        ms = min(ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(steps))
        events = int(cnt[_abi.CNT_EVENTS].item()) / steps
        print(json.dumps({"workload": name, "electrons": n, "seed": seed, "precision": "fp64" if fp64 else "fp32",
                          "launch_shape": tr.tuned, "ms_per_launch": ms, "events_per_launch": events,
                          "events_per_s": events / (ms * 1e-3)}), flush=True)


def cli_rate():
    """doScatter.py (the reference's command line, MC/doScatter.py) end to end:
    parsing the input file, the parameter-keyed tables (built on the first
    run, read back on the second), chunked launches and the HDF5 store, timed
    by wall clock on the DS Cu 20 keV 70 deg configuration."""
    import contextlib
    import io
    import re
    import shutil
    import tempfile

    import doScatter

    base = ("E0 : 20000\nEmin : 1000\nmaterial : Cu\ns_tilt : 70\nelastic : Ruth_vanilla\nmode : DS\n"
            "Bethe_model : classical\ndiffMFP : no\ntol_E : 1e-3\ntol_W : 1e-3\nmaxScatt : off\nWc : 50\n"
            "seed : 1\n")
    runs = (("els", 1 << 22, "", 1 << 20), ("els+scats", 1 << 12, "type, E_loss, position, direction", 1 << 10))
    d = tempfile.mkdtemp(prefix="mcd_cli_")
    out = []
    try:
        for gen in ("yes", "no"):
            for label, n, scat, chunk in runs:
                f = os.path.join(d, label + ".file")
                with open(f, "w") as fh:
                    fh.write(base + f"num_el : {n}\ngen_tables : {gen}\n" +
                             (f"scatter_output : {scat}\n" if scat else ""))
                buf = io.StringIO()
                t = time.perf_counter()
                with contextlib.redirect_stdout(buf):
                    path = doScatter.main(["-i", f, "-o", os.path.join(d, "out"), "-c", str(chunk)])
                wall = time.perf_counter() - t
                events = int(re.findall(r"(\d+) scatter events", buf.getvalue())[-1])
                out.append({"output": label, "gen_tables": gen, "electrons": n, "chunk": chunk,
                            "scatter_output": scat or None, "events": events, "wall_s": wall,
                            "events_per_s": events / wall, "store_mib": os.path.getsize(path) / 2**20})
                os.remove(path)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps({"workload": "doScatter.py end to end, DS Cu 20 keV -> 1 keV, 70 deg, Wc 50 eV, tol 1e-3, "
                                  "seed 1, HDF5 store", "runs": out}), flush=True)


def valu_ceiling(reps=5):
    """Wavefront-instructions per second the SIMDs issue for the VALU
    instructions that dominate the transport's arithmetic and its Philox
    rounds, alone and in pairs, each in a kernel of known instruction count at
    8 wavefronts per SIMD (mc-discrete_amd/csrc/mcd_calib.hip).  Timed with HIP
    events on the launch stream; the best of `reps` launches.  A mix issuing
    more than one instruction per CU per clock dual-issues (gfx950 SIMDs can
    issue two VALU instructions in one quad-cycle)."""
    import ctypes as C

    import torch

    L = C.CDLL(os.path.join(ROOT, "mc-discrete_amd", "mcd", "libmcd_calib.so"))
    L.mcd_calib_launch.restype = C.c_int
    L.mcd_calib_launch.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.mcd_calib_insts_per_wave.restype = C.c_longlong
    L.mcd_calib_insts_per_wave.argtypes = [C.c_int, C.c_int]
    dev = torch.device("cuda", 0)
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    blocks = n_cu * 8  # 8 workgroups of 4 wavefronts per CU: 8 wavefronts per SIMD
    waves = blocks * 4
    out = torch.empty(blocks * 256 * 4, dtype=torch.uint8, device=dev)
    s = torch.cuda.current_stream(dev)
    res = {"workload": f"{blocks} workgroups x 256 threads ({waves} wavefronts, 8 per SIMD on {n_cu} CUs)",
           "cus": n_cu}
    kinds = ((0, "v_fma_f32"), (1, "v_mad_u64_u32"), (2, "v_bitop3_b32"), (3, "v_fma_f32+v_exp_f32"),
             (4, "v_fma_f32+v_mad_u64_u32"), (5, "v_fma_f32+v_bitop3_b32"), (6, "v_fma_f32+v_alignbit_b32"),
             (7, "v_rcp_f32"))
    for kind, name in kinds:
        iters = 512 if kind in (1, 4) else 2048
        per_wave = L.mcd_calib_insts_per_wave(kind, iters)
        best = float("inf")
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            rc = L.mcd_calib_launch(kind, blocks, iters, out.data_ptr(), s.cuda_stream)
            e1.record(s)
            torch.cuda.synchronize(dev)
            if rc != 0:
                raise RuntimeError(f"calibration kernel {name} failed to launch ({rc})")
            best = min(best, e0.elapsed_time(e1))
        rate = per_wave * waves / (best * 1e-3)
        res[name] = {"insts_per_wave": per_wave, "ms": best, "wave_insts_per_s": rate,
                     "per_cu_per_clock_at_2_4GHz": rate / (n_cu * 2.4e9)}
    fma = res["v_fma_f32"]["wave_insts_per_s"]
    res["fp32_fma_tflops"] = fma * 64 * 2 / 1e12
    print(json.dumps(res), flush=True)
    return res


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU (asked for {a.gpus}); electron ranges would shard across GPUs, "
              "but no multi-GPU run is launched", file=sys.stderr)
        sys.exit(2)
    if a.config_sweep:
        config_sweep(a.steps, tune=not a.fixed_shapes)
        return
    if a.valu_ceiling:
        valu_ceiling()
        return
    if a.cli:
        cli_rate()
        return
    if a.precision:
        from mcd import precision
        for name, c in precision.run_study(a.electrons):
            print(json.dumps({"workload": name, **c}), flush=True)
        return
    import torch

    from mcd import _abi, engine

    dev = torch.device("cuda", 0)
    tr = engine.Transport("Cu", 20000.0, 1000.0, Wc=50.0, s_tilt=70.0, mode="DS", tol_E=1e-3, tol_W=1e-3)
    n = a.electrons
    els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    # launch shape (block, waves/SIMD, guard placement) from the autotuner,
    # outside the timed region; an explicit --block/--waves-per-simd skips it
    tuned = None
    if a.fixed_shapes and not (a.block or a.waves_per_simd or a.guard):
        a.block, a.waves_per_simd, a.guard = 1024, 8, "service"
    if not (a.block or a.waves_per_simd or a.guard):
        tuned = tr.tune(n, seed=a.seed, use_cache=False)
    for _ in range(a.warmup):
        tr.launch(n, els, cnt, seed=a.seed, block=a.block, waves_per_simd=a.waves_per_simd, guard=a.guard)
    torch.cuda.synchronize(dev)
    cnt.zero_()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for k in range(a.steps):
        starts[k].record()
        tr.launch(n, els, cnt, seed=a.seed, block=a.block, waves_per_simd=a.waves_per_simd, guard=a.guard)
        ends[k].record()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    kern_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / a.steps
    if a.calibrate_copy_mib:
        src = torch.empty(a.calibrate_copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize(dev)
        dst.copy_(src)
        torch.cuda.synchronize(dev)
        del src, dst
    c = cnt.cpu().numpy().view("uint64")
    engine.check_guard(c, "timed headline launches")  # records of a launch with unsettled picks are invalid
    events = int(c[_abi.CNT_EVENTS])
    value = events / wall
    ev_per_launch = events / a.steps
    achieved = ev_per_launch * F_OPS_CU_DS / (kern_ms * 1e-3) / 1e12
    # per-launch HBM bytes and VALU instructions per event of this kernel and
    # workload, from its committed rocprofv3 counter runs
    prof = {}
    ppath = os.path.join(ROOT, "profiles", "transport_ds_cu20_counters.json")
    if os.path.exists(ppath):
        with open(ppath) as f:
            prof = json.load(f)
    traffic = prof.get("hbm_bytes_per_launch")
    issue = None
    if prof.get("valu_insts_per_event"):
        rate = prof["valu_insts_per_event"] * ev_per_launch / (kern_ms * 1e-3)
        issue = {"valu_insts_per_event": prof["valu_insts_per_event"], "wave_insts_per_s": rate,
                 "ceiling_wave_insts_per_s": prof["valu_ceiling_wave_insts_per_s"],
                 "frac": rate / prof["valu_ceiling_wave_insts_per_s"],
                 "source": "VALU instructions per event: profiles/transport_ds_cu20_counters.json (rocprofv3 "
                           "SQ_INSTS_VALU); ceiling: v_fma_f32 issue rate measured by --valu-ceiling"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (pencil beam, Philox seed 1)",
        "config": {"workload": "DS transport, Cu bulk 20 keV -> 1 keV, s_tilt 70 deg, Wc 50 eV, "
                               "tol_E = tol_W = 1e-3, 2^22 electrons per step, seed 1",
                   "electrons_per_step": n, "parallelism": "single GPU"},
        "electrons_per_s": n * a.steps / wall,
        "events_per_electron": ev_per_launch / n,
        "backscatter_yield": float(c[_abi.CNT_BKS]) / (n * a.steps),
        "fp64_guard_picks_per_event": float(c[_abi.CNT_FP64_PICKS]) / max(events, 1),
        "kernel_ms": kern_ms,
        # the transport is bound by VALU issue (no MFMA, 0.06 ms of HBM time per
        # launch): algorithmic FP32 operations against the published FP32 peak,
        # and the measured issue rate against the measured VALU issue ceiling
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": "FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated "
                                       "(profiles/transport_ds_cu20_counters.json)",
                     "pipe": "VALU (FP32 and integer)", "ops_per_event": F_OPS_CU_DS, "issue": issue},
        "cpu_baseline": None,
        "launch_shape": ({"block": tuned["block"], "waves_per_simd": tuned["waves_per_simd"], "guard": tuned["guard"],
                          "refill_period": tuned.get("refill_period", 1),
                          "autotuned": True, "candidates_ms": tuned["ms"]} if tuned is not None else
                         {"block": a.block, "waves_per_simd": a.waves_per_simd, "guard": a.guard, "autotuned": False}),
    }
    if not a.no_two_stream:
        out["two_stream"] = two_stream(tr, n, a.steps, a.seed)
    if not a.no_backward:
        bw = backward_ratio(a.bwd_electrons, tune=not a.fixed_shapes)
        out["backward"] = bw
        out["bwd_fwd_ratio"] = bw["ratio_max"]
    if not a.no_batched:
        out["batched"] = batched_ratio(tune=not a.fixed_shapes)
    if not a.no_cpu_baseline:
        c_base, b1 = cpu_baselines(tr, a.cpu_seconds, a.seed)
        out["cpu_baseline"] = c_base
        out["cpu_baseline_reference_path"] = b1
        out["speedup_vs_cpu_baseline"] = value / c_base["value"]
        out["speedup_vs_reference_path"] = value / b1["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Launch-shape autotuners for single-configuration launches (mcd_forward) and
batched runs (mcd_batch_*).

Single configurations (``tune_transport``): ``block`` (threads per workgroup),
``waves_per_simd`` (the register budget: the resident wavefronts per SIMD it
must allow), ``guard``, where a DS step whose channel pick falls in the
guard band gets its FP64 pick: inline in the transport code, deferred to a
second kernel whose first phase carries no FP64 code, or served by wavefront 0
of each workgroup through LDS (one kernel whose transport code carries no
FP64), and ``refill_period``, the compaction period: a wavefront refills its
idle lanes with new electrons every that many steps (fewer refill passes,
lanes idle a little longer).  Records are identical for every shape.

Batched runs (``tune_batch``):

The batched kernel has four shape choices that trade against each other:

* ``pick_nodes`` — the channel-pick table layout: nodes per configuration in
  the wavefront's LDS slice.  More nodes give a narrower FP64 guard band and a
  finer mean-free-path interpolation but a larger slice, so fewer workgroups
  fit a CU; 0 sizes each table to the default 8 KB slice.
* ``block`` — threads per workgroup (64, 128 or 256).
* ``guard`` — where DS items take the FP64 guard pick: ``"deferred"`` (a
  first phase without FP64 code parks those electrons, a second resumes them),
  ``"inline"`` (one phase, more registers, fewer resident wavefronts) or
  ``"service"`` (one phase; wavefront 0 of each workgroup settles the picks of
  the others, which carry no FP64 code).
* ``refill_period`` — the compaction period, as for single configurations
  (idle lanes found by ballot and numbered by a prefix count).

``tune_batch`` times every candidate on the actual batch (graph replay, best
of ``reps``), keeps the fastest and caches the choice by a key of the
configurations and electron counts (JSON file under ``$MCD_TUNE_CACHE`` or
``~/.cache/mcd``).  Different node counts interpolate the mean free path
differently, so a run's records are reproducible for a given choice, not
across choices.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

CANDIDATES = (
    dict(block=1024, pick_nodes=128, guard="service"),
    dict(block=128, pick_nodes=128, guard="inline"),
    dict(block=256, pick_nodes=128, guard="inline"),
    dict(block=256, pick_nodes=0, guard="inline"),
    dict(block=256, pick_nodes=256, guard="inline"),
    dict(block=256, pick_nodes=128, guard="inline", refill_period=2),
    dict(block=256, pick_nodes=128, guard="inline", refill_period=4),
    dict(block=128, pick_nodes=128, guard="deferred"),
    dict(block=256, pick_nodes=128, guard="deferred"),
    dict(block=256, pick_nodes=0, guard="deferred"),
)


def _name(c):
    return (f"block={c['block']},pick_nodes={c['pick_nodes']},guard={c.get('guard') or 'default'},"
            f"refill_period={c.get('refill_period', 1)}")


TRANSPORT_CANDIDATES = (
    dict(block=256, waves_per_simd=4, guard="inline", refill_period=1),
    dict(block=128, waves_per_simd=4, guard="inline", refill_period=1),
    dict(block=256, waves_per_simd=5, guard="inline", refill_period=1),
    dict(block=256, waves_per_simd=4, guard="inline", refill_period=2),
    dict(block=256, waves_per_simd=0, guard="deferred", refill_period=1),
    dict(block=512, waves_per_simd=8, guard="service", refill_period=1),
    dict(block=1024, waves_per_simd=8, guard="service", refill_period=1),
    dict(block=512, waves_per_simd=7, guard="service", refill_period=1),
    dict(block=1024, waves_per_simd=8, guard="service", refill_period=2),
    dict(block=1024, waves_per_simd=8, guard="service", refill_period=4),
)


def _cache_path(name="batch_tune.json"):
    root = os.environ.get("MCD_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "mcd")
    return os.path.join(root, name)


def batch_key(items) -> str:
    """Key of a batch: every configuration field that shapes the work, and n."""
    h = hashlib.sha1()
    for tr, n in items:
        c = tr.cfg
        h.update(repr((c.mat.Z, round(c.mat.density, 9), c.E0, c.Emin, c.Wc, c.s_tilt, c.mode, c.bethe_model,
                       c.thickness, c.max_scatt, round(c.J, 9), int(n))).encode())
    return h.hexdigest()


def transport_key(tr, n) -> str:
    """Key of one configuration's launch: the fields that shape its work, and n."""
    return "t" + batch_key([(tr, n)])


def _tname(c):
    return (f"block={c['block']},waves_per_simd={c['waves_per_simd']},guard={c['guard']},"
            f"refill_period={c.get('refill_period', 1)}")


def _time_transport(tr, n, cand, seed, reps, device):
    import torch

    from . import _abi
    dev = torch.device("cuda", device)
    els = torch.empty(n * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    kw = dict(block=cand["block"], waves_per_simd=cand["waves_per_simd"], guard=cand["guard"],
              refill_period=cand.get("refill_period", 1))
    saved, tr.tuned = tr.tuned, None
    try:
        tr.launch(n, els, cnt, seed=seed, **kw)
        torch.cuda.synchronize(dev)
        best = float("inf")
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            tr.launch(n, els, cnt, seed=seed, **kw)
            b.record()
            torch.cuda.synchronize(dev)
            best = min(best, a.elapsed_time(b))
    finally:
        tr.tuned = saved
    return best


def _store(path, key, value):
    data = _load(path)
    data[key] = value
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + f".{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)


def tune_transport(tr, n=1 << 22, seed=1, candidates=TRANSPORT_CANDIDATES, reps=2, device=0, use_cache=True,
                   cache_path=None):
    """Pick the fastest launch shape for n electrons of Transport `tr` (FP32).

    Returns {"block", "waves_per_simd", "guard", "ms": {candidate: ms}, "cached": bool}."""
    path = cache_path or _cache_path("transport_tune.json")
    # the key names the candidate set too, so a cached choice is one of today's
    key = transport_key(tr, n) + hashlib.sha1(",".join(_tname(c) for c in candidates).encode()).hexdigest()[:8]
    from . import _abi
    if tr.mode != _abi.MODE_DS:  # continuous mode has no guard pick to place
        candidates = tuple(c for c in candidates if c["guard"] == "inline")
    if use_cache:
        hit = _load(path).get(key)
        if hit and any(_tname(hit) == _tname(c) for c in candidates):
            return dict(hit, cached=True)
    timings = {_tname(c): _time_transport(tr, n, c, seed, reps, device) for c in candidates}
    best = min(candidates, key=lambda c: timings[_tname(c)])
    out = dict(block=best["block"], waves_per_simd=best["waves_per_simd"], guard=best["guard"],
               refill_period=best.get("refill_period", 1), ms=timings)
    if use_cache:
        _store(path, key, out)
    return dict(out, cached=False)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _time_candidate(items, cand, seed, reps, device):
    import torch

    from .engine import Batch

    b = Batch(items, pick_nodes=cand["pick_nodes"], device=device)
    kw = dict(block=cand["block"], guard=cand.get("guard"), refill_period=cand.get("refill_period", 1))
    b.launch(seed=seed, use_graph=True, **kw)
    torch.cuda.synchronize(device)
    best = float("inf")
    for _ in range(reps):
        b.counters.zero_()
        torch.cuda.synchronize(device)
        t = time.perf_counter()
        b.launch(seed=seed, use_graph=True, **kw)
        torch.cuda.synchronize(device)
        best = min(best, time.perf_counter() - t)
    del b
    return best


def tune_batch(items, seed=3, candidates=CANDIDATES, reps=2, device=0, use_cache=True, cache_path=None):
    """Pick the fastest launch shape for `items` (list of (Transport, n)).

    Returns {"block", "pick_nodes", "guard", "refill_period", "ms": {candidate: ms}, "cached": bool}."""
    path = cache_path or _cache_path()
    # the key names the candidate set too, so a cached choice is one of today's
    key = batch_key(items) + hashlib.sha1(",".join(_name(c) for c in candidates).encode()).hexdigest()[:8]
    if use_cache:
        hit = _load(path).get(key)
        if hit and any(_name(hit) == _name(c) for c in candidates):
            return dict(hit, cached=True)
    timings = {}
    for cand in candidates:
        timings[_name(cand)] = 1e3 * _time_candidate(items, cand, seed, reps, device)
    best = min(candidates, key=lambda c: timings[_name(c)])
    out = dict(block=best["block"], pick_nodes=best["pick_nodes"], guard=best.get("guard"),
               refill_period=best.get("refill_period", 1), ms=timings)
    if use_cache:
        _store(path, key, out)
    return dict(out, cached=False)


__all__ = ["CANDIDATES", "TRANSPORT_CANDIDATES", "batch_key", "transport_key", "tune_batch", "tune_transport"]

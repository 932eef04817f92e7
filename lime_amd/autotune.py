"""Launch-shape autotuner for the Lindblad kernels.

The library exposes, per descriptor, a small set of launch-shape variants
(``lime_lindblad_variants``): for the resident kernel the (waves, tiles-per-wave) shapes of
each padded size, for the tiled kernels the 32x32 or 64x64 output tile.  ``tune`` times
every variant on the caller's own batch for a few RK4 steps (after one warm-up launch,
HIP events on the launch stream), keeps the fastest, and caches the choice per
(n, dense operators, precision, batch bucket, observables) in a JSON file
(``LIME_TUNE_CACHE``, else ``$XDG_CACHE_HOME/lime_amd/tune.json`` or
``~/.cache/lime_amd/tune.json``; never inside the package).
"""
import json
import os
import sys
import threading

import torch

_lock = threading.Lock()
_CACHE_PATH = os.environ.get("LIME_TUNE_CACHE") or os.path.join(
    os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache"),
    "lime_amd", "tune.json")
_cache = None


def _load():
    global _cache
    if _cache is None:
        try:
            with open(_CACHE_PATH) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _save():
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        tmp = _CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
        os.replace(tmp, _CACHE_PATH)
    except OSError:
        pass


def _bucket(batch):
    b = 1
    while b < batch:
        b *= 2
    return b


def key_of(prop, n_eops=0):
    return (f"n{prop.n}:kd{prop.n_dense}:{prop.precision}:b{_bucket(prop.batch)}"
            f":e{int(n_eops > 0)}:t{int(prop.force_tiled)}")


def tune(prop, rho, dt, steps=4, n_eops=0, force=False, verbose=False, eops_t=None, eop_group=None):
    """Pick the fastest launch variant for ``prop`` on operators shaped like ``rho``.

    ``rho`` is not modified (a scratch copy is advanced).  With ``eops_t`` (the transposed
    observables, as ``LindbladPropagator.step`` takes them) every timed step also records
    the observables, as a recording run does.  Sets ``prop.variant`` and returns
    (variant, {variant: ms}).
    """
    if eops_t is not None:
        n_eops = int(eops_t.shape[-3])
    with _lock:
        cache = _load()
        k = key_of(prop, n_eops)
        if not force and k in cache:
            prop.variant = int(cache[k]["variant"])
            return prop.variant, cache[k].get("ms", {})
    nv = prop.variants(n_eops=n_eops)
    if nv <= 1:
        prop.variant = 0
        return 0, {}
    scratch = rho.clone()
    times = {}
    stream = torch.cuda.current_stream(prop.dev)
    kw = {}
    if eops_t is not None:
        obs = torch.zeros((rho.shape[0], steps + 1, n_eops), dtype=torch.complex128, device=rho.device)
        kw = dict(eops_t=eops_t, eop_group=eop_group, obs=obs, nout=1, rec_capacity=steps + 1)
    for v in range(nv):
        prop.variant = v
        try:
            prop.step(scratch, dt, 1, **kw)                  # warm-up (and attribute set-up)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            prop.step(scratch, dt, steps, **kw)
            e1.record(stream)
            torch.cuda.synchronize(prop.dev)
            times[v] = e0.elapsed_time(e1) / steps
        except RuntimeError as err:                    # a variant this size cannot launch
            if verbose:
                print(f"[tune] variant {v} skipped: {err}", file=sys.stderr)
    best = min(times, key=times.get)
    prop.variant = best
    if verbose:
        print(f"[tune] {key_of(prop, n_eops)}: " + ", ".join(f"v{v}={t:.3f}ms" for v, t in sorted(times.items()))
              + f" -> v{best}", file=sys.stderr)
    with _lock:
        _load()[key_of(prop, n_eops)] = {"variant": best, "ms": {str(v): t for v, t in times.items()}}
        _save()
    return best, times


def tune_ragged(rb, dt, steps=2, force=False, verbose=False):
    """Pick the fastest persistent launch shape of a ``RaggedLindblad`` batch (states are
    restored afterwards); cached per (instance count, sum of n^3 (1 + K_d))."""
    key = f"ragged:c{rb.count}:w{int(rb.work_units)}"
    with _lock:
        cache = _load()
        if not force and key in cache:
            rb.variant = int(cache[key]["variant"])
            return rb.variant, cache[key].get("ms", {})
    saved = rb.rho.clone()
    times = {}
    stream = torch.cuda.current_stream(rb.dev)
    for v in range(rb.variants()):
        rb.variant = v
        rb.step(dt, 1)                                   # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rb.step(dt, steps)
        e1.record(stream)
        torch.cuda.synchronize(rb.dev)
        times[v] = e0.elapsed_time(e1) / steps
    rb.rho.copy_(saved)
    best = min(times, key=times.get)
    rb.variant = best
    if verbose:
        print(f"[tune] {key}: " + ", ".join(f"v{v}={t:.3f}ms" for v, t in sorted(times.items()))
              + f" -> v{best}", file=sys.stderr)
    with _lock:
        _load()[key] = {"variant": best, "ms": {str(v): t for v, t in times.items()}}
        _save()
    return best, times

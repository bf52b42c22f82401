"""Launch-shape autotuner for the pixel kernel.

Candidates are (pixels per lane K, waves per workgroup W, persistent or
grid launch, workgroups per CU for the persistent grid).  Each candidate is
timed with HIP events on the evaluation stream over a few repetitions of a
representative batch; the fastest is stored in a JSON cache keyed by device
architecture and workload kind, which ``load_or_default`` reads at start-up.
Launches within ``TIE`` (relative) of the fastest count as equal, and the
simplest of them wins: a grid launch before a persistent one (the persistent
kernel holds more registers and spills more), then fewer waves per
workgroup, then fewer workgroups per CU.  Results
do not depend on the persistent/grid choice or on W (same tiles, same fixed
reduction order); K changes the tiling and therefore only the rounding of
the FP32 partial sums.
"""

import itertools
import json
import os

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")
TIE = 0.01
DEFAULT = {"batch": dict(pix_per_lane=8, waves_per_group=1, persistent=0, groups_per_cu=8,
                         cull=1),
           "patch": dict(pix_per_lane=4, waves_per_group=4, persistent=0, groups_per_cu=8,
                         cull=1)}

__all__ = ["candidates", "pick", "tune", "load_or_default", "save"]


def candidates(kind="batch"):
    out = []
    for K, W in itertools.product((1, 2, 4, 8), (1, 2, 4)):
        if kind == "batch":
            for G in (4, 8, 16):
                out.append(dict(pix_per_lane=K, waves_per_group=W, persistent=1,
                                groups_per_cu=G, cull=1))
        out.append(dict(pix_per_lane=K, waves_per_group=W, persistent=0, groups_per_cu=8,
                        cull=1))
    return out


def _device_name():
    """Cache key: the GPU architecture (e.g. gfx950), which fixes the best
    launch shape; marketing names differ between tools on one device."""
    try:
        import torch
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:
        return "unknown"


def pick(table, tie=TIE):
    """The launch to keep from ``table`` = [(ms, launch), ...]: the simplest
    launch within ``tie`` of the fastest (grid before persistent, then fewer
    waves per workgroup, then fewer workgroups per CU)."""
    best = min(ms for ms, _ in table)
    close = [c for ms, c in table if ms <= best * (1.0 + tie)]
    return min(close, key=lambda c: (c["persistent"], c["waves_per_group"],
                                      c["groups_per_cu"]))


def tune(batch, run, cands=None, reps=3, log=None):
    """Time ``run()`` (one evaluation on torch's current stream) for every
    candidate launch of ``batch``; return (best, table)."""
    import torch
    cands = cands or candidates("batch")
    table = []
    for c in cands:
        batch.set_launch(**c)
        run()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            run()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        table.append((ms, c))
        if log:
            log(f"  {c}: {ms:.3f} ms")
    table.sort(key=lambda t: t[0])
    best = pick(table)
    batch.set_launch(**best)
    return best, table


def save(kind, launch, ms=None):
    data = {}
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            data = json.load(f)
    data.setdefault(_device_name(), {})[kind] = dict(launch=launch, ms=ms)
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def load_or_default(kind="batch"):
    """Cached best launch for this device and workload kind, else default."""
    if os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                data = json.load(f)
            entry = data.get(_device_name(), {}).get(kind)
            if entry:
                return dict(entry["launch"])
        except (OSError, ValueError, KeyError):
            pass
    return dict(DEFAULT[kind])

"""Autotuner of the extension kernels' launch table.

Sweeps, on the GPU, the knobs of the tuning table (include/bsx.h,
bsx_set_tuning): descriptors per block of the plain group kernels and of the
frame group kernels, and the group shape of the two query-length bins that have
alternatives (65-96 and 97-128 bases: 16 lanes x 6/8 columns or 32 lanes x 3/4
columns).  Coordinate descent from the current table: each knob in turn takes
the value with the lowest step time (plain form for the plain knobs, the
24-column frame form with plain fallback for the frame knob), every candidate's
outputs checked equal to the default table's.  The winner is cached in
bsbolt_amd/tune.json under the device's key ("<gfx arch>/<CU count>"), which
bsx_create loads.

    python -m bsbolt_amd.tune [--workload cfg1] [--scale 0.5] [--steps 5]
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tune.json")
KNOBS = (("tpb_plain", 0, (16, 32, 64, 128)), ("shape_var_65_96", 2, (0, 1)),
         ("shape_var_97_128", 3, (0, 1)), ("tpb_narrow", 1, (16, 32, 64)))


def _time(eng, ptrs, steps, reps=3):
    best = float("inf")
    for _ in range(reps):
        eng.extend_device(ptrs)
        eng.sync()
        eng.timer_start()
        for _ in range(steps):
            eng.extend_device(ptrs)
        best = min(best, eng.timer_stop() / steps)
    return best


def tune(eng, batch, steps=5, log=print):
    """Returns (table, record) for the resident batch."""
    from . import bsx
    from .tasks import RESULT_FIELDS
    n = len(batch)
    keep, ptrs = eng.tasks_to_device(batch)
    outs = {k: eng.empty((n,), np.int32) for k in RESULT_FIELDS}
    ptrs.update({k: v.ptr for k, v in outs.items()})
    ptrs["n"] = n
    eng.set_tuning(None)
    eng.extend_device(ptrs)
    eng.sync()
    ref = {k: v.numpy() for k, v in outs.items()}
    default = eng.get_tuning()[0]
    table = list(default)
    record = {"default_ms": {}, "sweep": []}
    for name, idx, values in KNOBS:
        frame = name == "tpb_narrow"
        if frame:
            eng.set_band_policy(bsx.BAND_FIXED, 0)
        try:
            times = {}
            for v in values:
                t = list(table)
                t[idx] = v
                eng.set_tuning(t)
                ms = _time(eng, ptrs, steps)
                for k in RESULT_FIELDS:
                    if not np.array_equal(outs[k].numpy(), ref[k]):
                        raise RuntimeError(f"tuning {name}={v} changed output {k}")
                times[v] = ms
                log(f"  {name}={v}: {ms:.3f} ms/step")
            table[idx] = min(times, key=times.get)
            record["sweep"].append({"knob": name, "ms": {str(k): v for k, v in times.items()},
                                    "chosen": table[idx]})
        finally:
            eng.set_band_policy(bsx.BAND_PLAIN)
    eng.set_tuning(default)
    record["default_ms"]["plain"] = _time(eng, ptrs, steps)
    eng.set_tuning(table)
    record["tuned_ms"] = {"plain": _time(eng, ptrs, steps)}
    del keep, outs
    return table, record


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cfg1")
    ap.add_argument("--scale", type=float, default=0.5)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--out", default=CACHE)
    args = ap.parse_args()
    from . import bsx
    from .workload import CONFIGS, Workload
    eng = bsx.Engine(0)
    wl = Workload(CONFIGS[args.workload].scaled(args.scale))
    batch = wl.build_ext_batch(lambda t: eng.extend(t, cells=False))[0]
    wl.close()
    key = eng.get_tuning()[2]
    print(f"tuning {key} on {args.workload} x {args.scale}: {len(batch)} tasks", flush=True)
    table, rec = tune(eng, batch, args.steps, log=lambda m: print(m, flush=True))
    cache = json.load(open(args.out)) if os.path.exists(args.out) else {}
    cache[key] = {"bsx_table": table,
                  "workload": f"{args.workload} x {args.scale} ({len(batch)} tasks)", **rec}
    with open(args.out, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    print(f"{key}: {table} (default {rec['default_ms']['plain']:.3f} ms, tuned {rec['tuned_ms']['plain']:.3f} ms) "
          f"-> {args.out}", flush=True)
    eng.close()


if __name__ == "__main__":
    sys.exit(main())

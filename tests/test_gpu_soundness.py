"""Learned-band soundness at scale: on every full-size workload batch (150 bp
PE, its disjoint training simulation, 250 bp indel-heavy, RRBS; > 10M
extensions together) the learned-band form's outputs equal the plain form's,
field by field including the cell count, and the plain form equals the CPU
oracle (pinned to the compiled reference by the golden fixtures)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off", "cells")
TOTAL = {}


@pytest.mark.parametrize("cfg", ["cfg1", "cfg2", "cfg3", "cfg4"])
def test_learned_equals_plain_equals_oracle(engine, cfg):
    from bsbolt_amd import band, bsx
    from bsbolt_amd.workload import CONFIGS, Workload
    w = Workload(CONFIGS[cfg])
    try:
        b, _, _, _ = w.build_ext_batch(lambda t: engine.extend(t, cells=False))
    finally:
        w.close()
    plain = engine.extend(b)
    engine.set_band_model(band.load_model())
    try:
        engine.set_band_policy(bsx.BAND_LEARNED)
        learned = engine.extend(b)
        st = engine.band_stats()
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)
    for k in FIELDS:
        bad = np.nonzero(learned[k] != plain[k])[0]
        assert len(bad) == 0, f"{cfg}: learned {k} differs at {len(bad)} tasks, first {bad[:5]}"
    want, _ = O.extend2_batch(b, nthreads=min(16, os.cpu_count() or 1))
    for k in FIELDS:
        bad = np.nonzero(plain[k] != want[k])[0]
        assert len(bad) == 0, f"{cfg}: plain {k} differs from the oracle at {len(bad)} tasks"
    narrow = int(st[:4, 0].sum())
    TOTAL[cfg] = len(b)
    print(f"{cfg}: {len(b)} extensions, {narrow} routed narrow, {int(st[:4, 1].sum())} rejected by the "
          f"frame check, 0 mismatches; running total {sum(TOTAL.values())}")

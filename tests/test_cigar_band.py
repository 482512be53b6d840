"""The CIGAR pass's band helpers against the oracle's restatement of
bwa_gen_cigar2's band (bwa.c:231-241) and closed forms of infer_bw
(bwamem.c:796-803)."""
import numpy as np

from oracle import oracle as O


def test_cigar_band_matches_oracle():
    from bsbolt_amd.workload import cigar_band
    rng = np.random.default_rng(0)
    for _ in range(3000):
        lq = int(rng.integers(1, 400))
        rl = int(lq + rng.integers(-30, 31))
        if rl < 1:
            continue
        w_ = int(rng.integers(0, 500))
        assert int(cigar_band(w_, lq, rl)) == O.global_band(lq, rl, w_)


def test_infer_bw_closed_forms():
    from bsbolt_amd.workload import infer_bw
    # equal lengths and a score within two gap costs of perfect: no band
    assert infer_bw(100, 100, 100 - 11, 1, 6, 1) == 0
    assert infer_bw(100, 100, 100 - 12, 1, 6, 1) == int((100 - 88 - 6) / 1 + 2)
    # unequal lengths: at least the length difference
    assert infer_bw(100, 120, 100, 1, 6, 1) == 20
    assert infer_bw(50, 40, 10, 1, 6, 1) == int((40 - 10 - 6) / 1 + 2)
    # vectorised
    v = infer_bw(np.array([100, 100]), np.array([100, 120]), np.array([89, 100]), 1, 6, 1)
    assert list(v) == [0, 20]

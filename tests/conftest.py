import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libbsx.so)")


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return {k: z[k] for k in z.files}


def golden_ext(name):
    from bsbolt_amd.tasks import ExtTasks, EXT_FIELDS
    z = load_golden(name)
    t = ExtTasks(*[z[k] for k in EXT_FIELDS])
    sc = dict(mat=z["mat"], o_del=int(z["o_del"]), e_del=int(z["e_del"]), o_ins=int(z["o_ins"]),
              e_ins=int(z["e_ins"]), zdrop=int(z["zdrop"]))
    ref = {k[4:]: z[k] for k in z if k.startswith("ref_")}
    return t, sc, ref


def golden_glb(name):
    from bsbolt_amd.tasks import GlbTasks, GLB_FIELDS
    z = load_golden(name)
    t = GlbTasks(*[z[k] for k in GLB_FIELDS])
    sc = dict(mat=z["mat"], o_del=int(z["o_del"]), e_del=int(z["e_del"]), o_ins=int(z["o_ins"]),
              e_ins=int(z["e_ins"]))
    return t, sc, z["ref_score"], z["ref_n_cigar"], z["ref_cigar"]


EXT_GOLDEN = ["ext_random_bsbolt.npz", "ext_edge_bsbolt.npz", "ext_brute_bsbolt.npz", "ext_random_nozdrop.npz",
              "ext_random_asym.npz", "ext_random_matrix.npz"]
GLB_GOLDEN = ["glb_random_bsbolt.npz", "glb_random_asym.npz"]


@pytest.fixture(scope="session")
def engine():
    from bsbolt_amd.bsx import Engine
    e = Engine(0)
    yield e
    e.close()


def workload_tasks(cfg, scale):
    """Extension batch of a scaled workload, its rounds run on the CPU oracle."""
    from bsbolt_amd.workload import Workload, CONFIGS
    from oracle import oracle as O
    w = Workload(CONFIGS[cfg].scaled(scale))
    try:
        b, res, side, rnd = w.build_ext_batch(lambda t: O.extend2_batch(t, nthreads=8)[0])
    finally:
        w.close()
    return b


@pytest.fixture(scope="session")
def wl_small():
    """~5k cfg0-style tasks (100 bp SE) and ~5.5k cfg3-style tasks (250 bp, indels)."""
    return workload_tasks("cfg0", 0.15), workload_tasks("cfg3", 0.004)

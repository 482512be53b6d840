"""bsbolt_amd: MI355X-native seed-extend engine for BSBolt's bisulfite aligner.

The product is libbsx.so (HIP kernels + C ABI, include/bsx.h); this package is
its Python host: ctypes bindings (bsx), task batches (tasks), the workload
generator, and the learned band predictor.
"""
from .tasks import ExtTasks, GlbTasks, random_ext_tasks  # noqa: F401

__all__ = ["ExtTasks", "GlbTasks", "random_ext_tasks"]

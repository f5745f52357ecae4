"""MI355X-native visual bundle adjustment for maplab's optv / optvi path.

Layers:
  abi        ctypes view of include/maplab_ba.h (the C ABI of libmaplab_ba.so)
  problem    structure-of-arrays visual maps and the seeded workload generator
  solver     DeviceProblem: the map resident in HBM, Evaluate and LM solve
  optimizer  maplab-shaped driver (ViProblemOptions, VIMapOptimizer)
"""
from . import abi, problem  # noqa: F401

__all__ = ["abi", "problem"]

"""MI355X-native spatial discretization of the 2-D compressible Euler equations in ALE form.

Drop-in for PyEulerALE's ``SpatialDiscretization`` (src/py_euler_ale/core.py) over PyTorch-ROCm
tensors, backed by hand-written gfx950 HIP kernels in ``libpea_hip.so``.
"""

NUM_DIM = 2
NUM_VAR = 4
HEAT_RATIO = 1.4

__all__ = ["NUM_DIM", "NUM_VAR", "HEAT_RATIO"]

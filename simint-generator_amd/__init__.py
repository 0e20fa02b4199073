"""MI355X-native electron-repulsion-integral engine with simint's C API.

The product is libsimint.so (built by build.py): the simint C API, a HIP
runtime, and generated gfx950 class kernels.  The Python modules here are
the host-side mirror used by tests, the benchmark and the autotuner:
  simint     ctypes binding of the C API (shells, shell pairs, compute calls)
  workloads  synthetic shell sets of the benchmark configurations
  molfile    reader for simint's .mol test fixtures
  emitter/   the kernel emitter (recurrence graphs -> HIP sources)
"""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

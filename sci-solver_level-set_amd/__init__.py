"""MI355X-native level-set hot path (seeds, FIM redistancing, narrowband,
advection) behind a C-ABI library; this package holds the HIP/C++ sources,
the ctypes binding used by tests and benchmarks, and workload generators.

The directory name is not a Python identifier, so callers load it with
``load_package()`` from the repository root (see ``__graft_entry__.py``) and
import it as ``levelset_amd``.
"""

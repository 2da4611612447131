"""msa_amd -- MI355X-native all-pairs Needleman-Wunsch stage of Museqa.

The compute path is libmsa_hip.so (C ABI in include/msa_hip.h): hand-written
HIP kernels for gfx950 plus their host planner.  `msa_amd._native` binds it
with ctypes; `msa_amd.museqa` is the Python mirror of the reference's
museqa.sequence / museqa.database / museqa.pairwise modules over the C++ host
layer.
"""
__all__ = ["workloads"]

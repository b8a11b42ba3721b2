"""MI355X-native GENIE match-count top-k search.

The product is native: the HIP engine (``lib/libgenie_hip.so``, C ABI in
``include/genie_hip.h``) and the C++ host library that keeps GENIE's
table-build / query / top-k API (``lib/libgenie.so``, ``include/genie/``).
This package only loads them with ctypes and moves numpy arrays across; it
has no compute path of its own.  Importing it raises when the libraries are
missing: run ``python genie_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from ._native import (DIM_DTYPE, RANGE_DTYPE, GenieError, Index, Plan, Server, Table, device_synchronize,
                      knn_mt_admit, lib_paths, match_topk, mem_info, synth, tables_knn_mt, wire)

__all__ = ["DIM_DTYPE", "RANGE_DTYPE", "GenieError", "Index", "Plan", "Server", "Table", "device_synchronize",
           "knn_mt_admit", "lib_paths", "match_topk", "mem_info", "synth", "tables_knn_mt", "wire"]

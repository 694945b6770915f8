"""MI355X-native wavefront path tracer.

The drop-in boundary is the C ABI of libwfpt.so (include/wfpt.h); this package
loads it and offers a Python view of the same stage API for tests and the
benchmark.  Importing the package does not need a GPU; calling a stage does.
"""
from . import _capi
from ._capi import WfError, load, missing_symbols

__all__ = ["WfError", "load", "missing_symbols", "version"]


def version() -> str:
    return load().wf_version().decode()

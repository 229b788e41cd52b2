"""MI355X-native nlchains hot path: gfx950 kernels behind the C ABI of include/nlc.h.

The package holds the HIP sources (csrc/), the C++ nlchains CLI (host/), the
build recipe (build.py) and a ctypes binding (api.py).  The binding is loaded
on first use of any public name, and raises ImportError when
libnlchains_amd.so has not been built (there is no CPU fallback).
"""
import importlib


def __getattr__(name):
    if name.startswith("__"):
        raise AttributeError(name)
    api = importlib.import_module(".api", __name__)
    return getattr(api, name)

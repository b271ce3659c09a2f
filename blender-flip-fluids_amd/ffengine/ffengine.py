"""Library loader (counterpart of src/engine/ffengine/ffengine.py:49-134).

The engine library is built in-tree by csrc/Makefile into ../lib/.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.normpath(os.path.join(_HERE, os.pardir, "lib"))
LIB_PATH = os.path.join(LIB_DIR, "libffengine.so")


class FFEngineLib:
    def __init__(self):
        self._lib = None

    def _load_library(self):
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(
                "libffengine.so not found at %s: build it first "
                "(python -c 'import __graft_entry__ as g; g.build()')" % LIB_PATH)
        self._lib = ctypes.cdll.LoadLibrary(LIB_PATH)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if self._lib is None:
            self._load_library()
        return getattr(self._lib, name)

    @property
    def path(self):
        return LIB_PATH


ffengine = FFEngineLib()

"""build() of the repository entry point compiles (incrementally) and finds
the engine library it reports."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_build_entry():
    spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.build()
    from ffengine.ffengine import ffengine as lib
    assert os.path.isfile(lib.path)

"""ctypes calls into libffengine.so under the engine's error convention.

Every entry point except FluidSimulation_destroy and
CBindings_get_error_message ends with an int* flag that the library sets to
1 on success and to 0 on failure, with the reason left in one global message
buffer (cbindings.cpp:35-43).  A failed call raises
RuntimeError("<entry point> - <message>"), as the reference's Python layer
does (src/engine/ffengine/pybindings.py:26-59); the helper names are the ones
its callers use.
"""
import ctypes

from .ffengine import ffengine as lib

# scalar C types that a numpy scalar or a bool must be converted to before the call
_SCALARS = (ctypes.c_int, ctypes.c_uint, ctypes.c_long, ctypes.c_longlong, ctypes.c_size_t,
            ctypes.c_float, ctypes.c_double)

_get_message = lib.CBindings_get_error_message
_get_message.argtypes = []
_get_message.restype = ctypes.c_char_p


def engine_message():
    """The library's last error message."""
    raw = _get_message()
    return raw.decode("utf-8", "replace") if raw else ""


def check_success(flag, prefix):
    """Raises RuntimeError(prefix + the library's message) for a zero flag."""
    if not flag:
        raise RuntimeError(prefix + engine_message())


def init_lib_func(fn, argtypes, restype):
    """Declares an entry point's parameters (the trailing flag pointer
    included) and its result type; later declarations of the same entry
    point are ignored."""
    if fn.argtypes is None:
        fn.argtypes = list(argtypes)
        fn.restype = restype


def execute_lib_func(fn, params):
    """Calls a declared entry point with `params` and a fresh success flag,
    raises on failure and returns the call's result (None for void)."""
    declared = fn.argtypes or []
    call_args = [ctype(value) if ctype in _SCALARS and not isinstance(value, ctype) else value
                 for ctype, value in zip(declared, params)]
    call_args.extend(params[len(call_args):])
    flag = ctypes.c_int(0)
    result = fn(*call_args, ctypes.byref(flag))
    check_success(flag.value, fn.__name__ + " - ")
    return result

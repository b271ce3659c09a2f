"""ctypes call helpers (counterpart of src/engine/ffengine/pybindings.py:26-59).

Every engine call ends with an int* success flag; a zero flag raises
RuntimeError("<function name> - <engine message>").
"""
from ctypes import c_char_p, c_int, byref

from .ffengine import ffengine as lib


def check_success(success, errprefix):
    libfunc = lib.CBindings_get_error_message
    init_lib_func(libfunc, [], c_char_p)
    if not success:
        raise RuntimeError(errprefix + str(libfunc().decode("utf-8")))


def init_lib_func(libfunc, argtypes, restype):
    if libfunc.argtypes is None:
        libfunc.argtypes = argtypes
        libfunc.restype = restype


def execute_lib_func(libfunc, params):
    args = []
    for idx, arg in enumerate(params):
        try:
            cval = libfunc.argtypes[idx](arg)
        except Exception:
            cval = arg
        args.append(cval)
    success = c_int()
    args.append(byref(success))
    if libfunc.restype:
        funcresult = libfunc(*args)
        check_success(success.value, libfunc.__name__ + " - ")
        try:
            return libfunc.restype(funcresult).value
        except Exception:
            return funcresult
    libfunc(*args)
    check_success(success.value, libfunc.__name__ + " - ")
    return None

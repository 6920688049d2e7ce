"""C ABI checks that need no GPU: the library loads, exports every
function include/siminf_hip.h declares, and its host-only helpers answer
like the reference (model registry, error messages)."""
import ctypes
import os
import re

import numpy as np

import siminf_amd as si
from siminf_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "siminf_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(SimInf_hip_\w+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = _native.load()
    names = declared_functions()
    assert len(names) >= 13
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTED_SYMBOLS)


def test_struct_layout_matches_header(tmp_path):
    """Offsets of every SimInf_solver_args field as gcc lays out the
    header, against the ctypes mirror."""
    import subprocess
    fields = [f[0] for f in _native.SolverArgs._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "siminf_hip.h"\n'
                   "int main(void){printf(\"%zu\\n\", sizeof(SimInf_solver_args));" +
                   "".join('printf("%%zu\\n", offsetof(SimInf_solver_args, %s));' % f
                           for f in fields) + "return 0;}")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(_native.SolverArgs)
    for name, off in zip(fields, out[1:]):
        assert getattr(_native.SolverArgs, name).offset == off, name


def test_builtin_models_and_dims():
    lib = _native.load()
    dims = {"SIR": (3, 2, 0), "SEIR": (4, 3, 0), "SIS": (2, 2, 0),
            "SISe": (2, 2, 1), "SISe3": (6, 6, 1), "SIR_mparse": (3, 2, 0)}
    for name, d in dims.items():
        h = lib.SimInf_hip_model_builtin(name.encode())
        assert h
        out = [ctypes.c_int() for _ in range(3)]
        assert lib.SimInf_hip_model_dims(ctypes.c_void_p(h), *[ctypes.byref(o) for o in out]) == 0
        assert tuple(o.value for o in out) == d
    assert not lib.SimInf_hip_model_builtin(b"nope")


def test_error_messages_are_the_references():
    lib = _native.load()
    msg = lambda c: lib.SimInf_hip_error_message(c).decode()
    assert msg(-1) == "Negative state detected."
    assert msg(-8) == "Unable to sample individuals for event."
    assert msg(-11) == "Invalid rate detected (non-finite or < 0.0)."
    assert msg(-12) == "Invalid 'solver' value."
    assert msg(-19) == "Cannot run the 'aem' solver on a replicated model."


def test_events_sorted_like_SimInf_events():
    ev = {"event": [3, 1, 0, 1, 0], "time": [2, 1, 2, 1, 1], "node": [1, 2, 3, 4, 5],
          "dest": [2, 0, 0, 0, 0], "n": [1] * 5, "proportion": [0.0] * 5,
          "select": [1, 2, 1, 1, 1], "shift": [0] * 5}
    e = si.make_events(np.eye(3), None, ev)
    # order(time, event, select), stable (R/SimInf_events.R:398)
    assert list(e["node"]) == [5, 4, 2, 3, 1]


def test_invalid_solver_rejected_before_any_device_call():
    m = si.SIR(u0={"S": [9], "I": [1], "R": [0]}, tspan=[1.0, 2.0], beta=0.1, gamma=0.1)
    try:
        si.run(m, solver="nope")
    except si.SimInfError as e:
        assert e.code == -12 and str(e) == "Invalid 'solver' value."
    else:
        raise AssertionError("expected an error")

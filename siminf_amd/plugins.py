"""Device plugins for models the parser generates: HIP source from the
model's function bodies, compiled with hipcc into a shared object and
cached by a digest of the source (the reference compiles its generated C
with R CMD SHLIB and caches the DLL by a digest of the code:
R/run.R:39-98).

The device plugin instantiates the solver kernels (csrc/siminf_kernels.h)
for the generated model and exports

    const SimInf_hip_model *siminf_plugin_model(void);

which the library runs like a built-in model.  host_source() gives the
same functions as plain C (TRFun/PTSFun table) for CPU callers.
"""
import ctypes
import hashlib
import os
import subprocess
import threading

from . import build as _build

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
CACHE = os.path.join(HERE, "_plugins")

_loaded = {}


def _fun_lines(parsed):
    """pre_code, trFunN and ptsFun exactly as the C generator emits them
    (the R-facing run/registration parts dropped)."""
    from . import modelparser as mp
    return (mp.C_enum(parsed.compartments_names, parsed.ldata_names,
                      parsed.gdata_names, parsed.v0_names, parsed.use_enum) +
            mp.C_pre_code(parsed.pre_code) +
            mp.C_trFun(parsed.props, parsed.variables) +
            mp.C_ptsFun(parsed.pts_fun))


# The library holds sets of compartments, transitions and continuous
# states as 32-bit masks (SIMINF_MAX_SET in include/siminf_hip.h).
MAX_SET = 32


def device_source(parsed, name):
    NC, NT = parsed.S.shape
    ND = parsed.Nd
    if max(NC, NT, ND) > MAX_SET:
        raise ValueError("The GPU solver supports at most %d compartments, %d transitions "
                         "and %d continuous states." % (MAX_SET, MAX_SET, MAX_SET))
    NLD = parsed.Nld
    NGD = parsed.Ngdata
    S = [[int(parsed.S[c, k]) for c in range(NC)] for k in range(NT)]
    G = [[int(parsed.G[j, k] != 0) for j in range(NT)] for k in range(NT)]
    fmt = lambda rows: "{" + ", ".join("{" + ", ".join(str(x) for x in r) + "}"
                                       for r in rows) + "}"
    lines = ["/* Device plugin generated from the model's C code. */",
             '#include "siminf_kernels.h"',
             '#include "siminf_plugin_api.h"',
             "",
             "#pragma clang force_cuda_host_device begin",
             "#pragma clang attribute push (__attribute__((always_inline)), apply_to = function)"]
    lines += _fun_lines(parsed)
    lines += ["#pragma clang attribute pop",
              "#pragma clang force_cuda_host_device end", ""]
    lines += ["struct GeneratedModel {",
              "    static constexpr int NC = %d, NT = %d, ND = %d, NLD = %d, NGD = %d;"
              % (NC, NT, ND, NLD, NGD),
              "    static constexpr bool PTS_NOOP = %s;" % ("true" if parsed.pts_fun is None
                                                           else "false"),
              "    static constexpr int S[NT][NC] = %s;" % fmt(S),
              "    static constexpr int G[NT][NT] = %s;" % fmt(G),
              "    template <int J> SIMINF_MODEL_FN double rate(SIMINF_RATE_ARGS)",
              "    {"]
    for k in range(NT):
        kw = "if constexpr" if k == 0 else "else if constexpr"
        if k == NT - 1 and NT > 1:
            lines.append("        else return trFun%d(u, v, ld, gd, t);" % (k + 1))
        elif NT == 1:
            lines.append("        return trFun1(u, v, ld, gd, t);")
        else:
            lines.append("        %s (J == %d) return trFun%d(u, v, ld, gd, t);"
                         % (kw, k, k + 1))
    lines += ["    }",
              "    SIMINF_MODEL_FN int pts(SIMINF_PTS_ARGS)",
              "    {",
              "        return ptsFun(vn, u, v, ld, gd, node, t);",
              "    }",
              "};",
              "",
              'extern "C" const SimInf_hip_model *siminf_plugin_model(void)',
              "{",
              '    static const SimInf_hip_model m = SIMINF_MODEL_ENTRY("%s", GeneratedModel, %d);'
              % (name, NGD),
              "    return &m;",
              "}", ""]
    return "\n".join(lines)


def host_source(parsed):
    NT = parsed.S.shape[1]
    lines = ["/* Host plugin generated from the model's C code. */",
             '#include "siminf_plugin_api.h"', ""]
    lines += _fun_lines(parsed)
    lines += ["typedef double (*TRFun)(const int *, const double *, const double *,",
              "                      const double *, double);",
              "typedef int (*PTSFun)(double *, const int *, const double *,",
              "                      const double *, const double *, int, double);",
              "TRFun siminf_tr_fun[] = {%s};" % ", ".join("&trFun%d" % (k + 1)
                                                       for k in range(NT)),
              "PTSFun siminf_pts_fun = &ptsFun;", ""]
    return "\n".join(lines)


def _digest(text, flags):
    h = hashlib.sha256()
    h.update(text.encode())
    h.update(" ".join(flags).encode())
    for f in ("siminf_kernels.h", "siminf_device.h", "siminf_rand.h",
              "siminf_models.h", "siminf_plugin_api.h"):
        h.update(open(os.path.join(CSRC, f), "rb").read())
    h.update(open(os.path.join(INCLUDE, "siminf_hip.h"), "rb").read())
    return h.hexdigest()[:24]


# Compiles and loads are serialised: the serving loop's worker threads may
# ask for the same plugin at once (they share the process id the temporary
# file names carry).
_lock = threading.RLock()


def compile_device(parsed, name="mparse"):
    """Path of the device plugin for a parsed model, compiling it on first
    use (cached under siminf_amd/_plugins/ by a digest of the source)."""
    with _lock:
        return _compile_device(parsed, name)


def _compile_device(parsed, name):
    src = device_source(parsed, name)
    flags = list(_build.FLAGS) + ["-shared"]
    key = _digest(src, flags)
    os.makedirs(CACHE, exist_ok=True)
    so = os.path.join(CACHE, "dev_%s.so" % key)
    if not os.path.exists(so):
        path = os.path.join(CACHE, "dev_%s.hip" % key)
        with open(path, "w") as f:
            f.write(src)
        tmp = so + ".tmp%d" % os.getpid()
        cmd = [_build.HIPCC] + flags + ["-I", CSRC, "-I", INCLUDE, "-o", tmp, path]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for the generated model:\n" + r.stderr[-4000:])
        os.replace(tmp, so)
    return so


def compile_host(parsed):
    """Host (C) plugin of the same functions, for CPU callers."""
    with _lock:
        return _compile_host(parsed)


def _compile_host(parsed):
    src = host_source(parsed)
    flags = ["-O2", "-fPIC", "-shared", "-ffp-contract=off"]
    key = _digest(src, flags)
    os.makedirs(CACHE, exist_ok=True)
    so = os.path.join(CACHE, "host_%s.so" % key)
    if not os.path.exists(so):
        path = os.path.join(CACHE, "host_%s.c" % key)
        with open(path, "w") as f:
            f.write(src)
        tmp = so + ".tmp%d" % os.getpid()
        r = subprocess.run(["gcc"] + flags + ["-I", CSRC, "-o", tmp, path, "-lm"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("gcc failed for the generated model:\n" + r.stderr[-4000:])
        os.replace(tmp, so)
    return so


def clear_cache():
    """Remove every cached plugin source and object."""
    if os.path.isdir(CACHE):
        for f in os.listdir(CACHE):
            if f.startswith(("dev_", "host_")):
                os.remove(os.path.join(CACHE, f))


def load_device(so):
    """SimInf_hip_model* of a compiled device plugin."""
    with _lock:
        if so not in _loaded:
            lib = ctypes.CDLL(so)
            lib.siminf_plugin_model.restype = ctypes.c_void_p
            _loaded[so] = (lib, lib.siminf_plugin_model())
        return _loaded[so][1]


class GeneratedPlugin:
    """A parsed model's device plugin (compiled lazily)."""

    def __init__(self, parsed, name="mparse"):
        self.parsed = parsed
        self.name = name
        self._so = None

    def device_model(self):
        if self._so is None:
            self._so = compile_device(self.parsed, self.name)
        return load_device(self._so)

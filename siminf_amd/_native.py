"""ctypes binding of the C ABI in include/siminf_hip.h.

The library ``libsiminf_hip.so`` is built in-tree (``siminf_amd.build``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsiminf_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)


class SolverArgs(ctypes.Structure):
    """Mirror of SimInf_solver_args (src/SimInf_internal.h:44-200)."""

    _fields_ = [
        ("u0", c_int_p),
        ("v0", c_double_p),
        ("irG", c_int_p),
        ("jcG", c_int_p),
        ("irS", c_int_p),
        ("jcS", c_int_p),
        ("prS", c_int_p),
        ("tspan", c_double_p),
        ("tlen", ctypes.c_int),
        ("U", c_int_p),
        ("irU", c_int_p),
        ("jcU", c_int_p),
        ("prU", c_double_p),
        ("V", c_double_p),
        ("irV", c_int_p),
        ("jcV", c_int_p),
        ("prV", c_double_p),
        ("ldata", c_double_p),
        ("gdata", c_double_p),
        ("Nn", ctypes.c_int),
        ("Nc", ctypes.c_int),
        ("Nt", ctypes.c_int),
        ("Nd", ctypes.c_int),
        ("Nld", ctypes.c_int),
        ("Nrep", ctypes.c_int),
        ("irE", c_int_p),
        ("jcE", c_int_p),
        ("prE", c_double_p),
        ("N", c_int_p),
        ("len", ctypes.c_int),
        ("event", c_int_p),
        ("time", c_int_p),
        ("node", c_int_p),
        ("dest", c_int_p),
        ("n", c_int_p),
        ("proportion", c_double_p),
        ("select", c_int_p),
        ("shift", c_int_p),
        ("Nthread", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("crn", ctypes.c_int),
        ("rng_type", ctypes.c_int),
        ("seed", ctypes.c_ulong),
        ("tr_fun", ctypes.c_void_p),
        ("pts_fun", ctypes.c_void_p),
    ]


# Every extern "C" symbol declared in include/siminf_hip.h.
EXPORTED_SYMBOLS = (
    "SimInf_hip_model_builtin",
    "SimInf_hip_model_dims",
    "SimInf_hip_run_solver_ssm",
    "SimInf_hip_run_solver_mssm",
    "SimInf_hip_run_solver_mssm_crn",
    "SimInf_hip_run_solver_aem",
    "SimInf_hip_run",
    "SimInf_hip_error_message",
    "SimInf_hip_plan_create",
    "SimInf_hip_plan_run",
    "SimInf_hip_plan_download",
    "SimInf_hip_plan_stats",
    "SimInf_hip_plan_kernel_times",
    "SimInf_hip_plan_set_replicate_data",
    "SimInf_hip_abc_weights_noalloc",
    "SimInf_hip_release",
    "SimInf_hip_batch_create",
    "SimInf_hip_batch_run",
    "SimInf_hip_batch_set_graph",
    "SimInf_hip_batch_autotune",
    "SimInf_hip_batch_get_shape",
    "SimInf_hip_batch_time",
    "SimInf_hip_batch_destroy",
    "SimInf_hip_pfilter",
    "SimInf_hip_plan_prevalence",
    "SimInf_hip_plan_fuse_prevalence",
    "SimInf_hip_plan_trajectory",
    "SimInf_hip_plan_trajectory_v",
    "SimInf_hip_abc_weights",
    "SimInf_hip_plan_set_lanes_per_wave",
    "SimInf_hip_plan_set_fill_block",
    "SimInf_hip_plan_last_error",
    "SimInf_hip_plan_get_fill_block",
    "SimInf_hip_plan_set_herd_map",
    "SimInf_hip_plan_get_herd_map",
    "SimInf_hip_plan_autotune",
    "SimInf_hip_plan_destroy",
)

_lib = None


def load(path=None):
    """Load the HIP library (once) and declare the prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            "siminf_amd: the HIP library %s is not built; run "
            "siminf_amd.build.build() first" % path)
    lib = ctypes.CDLL(path)
    args_p = ctypes.POINTER(SolverArgs)
    vp = ctypes.c_void_p
    lib.SimInf_hip_model_builtin.restype = vp
    lib.SimInf_hip_model_builtin.argtypes = [ctypes.c_char_p]
    lib.SimInf_hip_model_dims.restype = ctypes.c_int
    lib.SimInf_hip_model_dims.argtypes = [vp, c_int_p, c_int_p, c_int_p]
    for name in ("SimInf_hip_run_solver_ssm", "SimInf_hip_run_solver_mssm",
                 "SimInf_hip_run_solver_mssm_crn", "SimInf_hip_run_solver_aem",
                 "SimInf_hip_run"):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [args_p, vp]
    lib.SimInf_hip_error_message.restype = ctypes.c_char_p
    lib.SimInf_hip_error_message.argtypes = [ctypes.c_int]
    lib.SimInf_hip_plan_create.restype = ctypes.c_int
    lib.SimInf_hip_plan_create.argtypes = [ctypes.POINTER(vp), args_p, vp]
    lib.SimInf_hip_plan_run.restype = ctypes.c_int
    lib.SimInf_hip_plan_run.argtypes = [vp, ctypes.c_ulong]
    lib.SimInf_hip_plan_download.restype = ctypes.c_int
    lib.SimInf_hip_plan_download.argtypes = [vp, c_int_p, c_double_p,
                                             c_double_p, c_double_p]
    lib.SimInf_hip_plan_stats.restype = ctypes.c_int
    lib.SimInf_hip_plan_stats.argtypes = [
        vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
        ctypes.POINTER(ctypes.c_double)]
    lib.SimInf_hip_plan_kernel_times.restype = ctypes.c_int
    lib.SimInf_hip_plan_kernel_times.argtypes = [
        vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
        ctypes.POINTER(ctypes.c_longlong)]
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.SimInf_hip_pfilter.restype = ctypes.c_int
    lib.SimInf_hip_pfilter.argtypes = [
        args_p, vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, ctypes.c_int, ip, ctypes.c_ulong, dp, dp, ip]
    lib.SimInf_hip_plan_prevalence.restype = ctypes.c_int
    lib.SimInf_hip_plan_prevalence.argtypes = [
        vp, ctypes.c_uint, ctypes.c_uint, ip, dp, ctypes.c_int, ctypes.c_int, ip,
        ctypes.c_int, dp]
    lib.SimInf_hip_plan_fuse_prevalence.restype = ctypes.c_int
    lib.SimInf_hip_plan_fuse_prevalence.argtypes = [vp, ctypes.c_uint, ctypes.c_uint]
    lib.SimInf_hip_plan_trajectory.restype = ctypes.c_int
    lib.SimInf_hip_plan_trajectory.argtypes = [vp, ctypes.c_uint, ip, ctypes.c_int, ip]
    lib.SimInf_hip_plan_trajectory_v.restype = ctypes.c_int
    lib.SimInf_hip_plan_trajectory_v.argtypes = [vp, ctypes.c_uint, ip, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_double)]
    for name in ("SimInf_hip_abc_weights", "SimInf_hip_abc_weights_noalloc"):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_char_p, dp, dp, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp,
                      dp]
    lib.SimInf_hip_plan_set_replicate_data.restype = ctypes.c_int
    lib.SimInf_hip_plan_set_replicate_data.argtypes = [vp, dp, ctypes.c_int, dp]
    lib.SimInf_hip_plan_set_lanes_per_wave.restype = ctypes.c_int
    lib.SimInf_hip_plan_set_lanes_per_wave.argtypes = [vp, ctypes.c_int]
    lib.SimInf_hip_plan_set_fill_block.restype = ctypes.c_int
    ip_ = ctypes.POINTER(ctypes.c_int)
    lib.SimInf_hip_plan_last_error.restype = ctypes.c_int
    lib.SimInf_hip_plan_last_error.argtypes = [vp, ip_, ip_, ip_, ip_,
                                                ctypes.POINTER(ctypes.c_longlong)]
    lib.SimInf_hip_plan_set_fill_block.argtypes = [vp, ctypes.c_int]
    lib.SimInf_hip_plan_get_fill_block.restype = ctypes.c_int
    lib.SimInf_hip_plan_get_fill_block.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    lib.SimInf_hip_plan_set_herd_map.restype = ctypes.c_int
    lib.SimInf_hip_plan_set_herd_map.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                                 ctypes.c_int]
    lib.SimInf_hip_plan_get_herd_map.restype = ctypes.c_int
    lib.SimInf_hip_plan_get_herd_map.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
    lib.SimInf_hip_plan_autotune.restype = ctypes.c_int
    lib.SimInf_hip_plan_autotune.argtypes = [vp, ctypes.c_ulong, ctypes.c_int, ip, dp]
    lib.SimInf_hip_plan_destroy.restype = None
    lib.SimInf_hip_plan_destroy.argtypes = [vp]
    lib.SimInf_hip_release.restype = ctypes.c_int
    lib.SimInf_hip_release.argtypes = []
    lib.SimInf_hip_batch_create.restype = ctypes.c_int
    lib.SimInf_hip_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp),
                                            ctypes.c_int]
    lib.SimInf_hip_batch_run.restype = ctypes.c_int
    lib.SimInf_hip_batch_run.argtypes = [vp, ctypes.POINTER(ctypes.c_ulong)]
    lib.SimInf_hip_batch_autotune.restype = ctypes.c_int
    lib.SimInf_hip_batch_autotune.argtypes = [vp, ctypes.c_ulong, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_double)]
    lib.SimInf_hip_batch_get_shape.restype = ctypes.c_int
    lib.SimInf_hip_batch_get_shape.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int)]
    lib.SimInf_hip_batch_set_graph.restype = ctypes.c_int
    lib.SimInf_hip_batch_set_graph.argtypes = [vp, ctypes.c_int]
    lib.SimInf_hip_batch_time.restype = ctypes.c_int
    lib.SimInf_hip_batch_time.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.SimInf_hip_batch_destroy.restype = None
    lib.SimInf_hip_batch_destroy.argtypes = [vp]
    if path == LIB_PATH:
        _lib = lib
    return lib


def ptr(a, ctype):
    """Pointer to the data of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class ArgsBundle:
    """A SolverArgs plus the numpy arrays it points into (kept alive)."""

    def __init__(self, model, solver=0, crn=0, seed=0, dense=True):
        m = model
        self.keep = []

        def ia(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).ravel(order="F"))
            self.keep.append(a)
            return a

        def da(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))
            self.keep.append(a)
            return a

        a = SolverArgs()
        self.u0 = ia(m.u0_full())
        a.u0 = ptr(self.u0, ctypes.c_int)
        self.v0 = da(m.v0_full())
        a.v0 = ptr(self.v0, ctypes.c_double)
        G, S, E = m.G.tocsc(), m.S.tocsc(), m.E.tocsc()
        a.irG = ptr(ia(G.indices), ctypes.c_int)
        a.jcG = ptr(ia(G.indptr), ctypes.c_int)
        a.irS = ptr(ia(S.indices), ctypes.c_int)
        a.jcS = ptr(ia(S.indptr), ctypes.c_int)
        a.prS = ptr(ia(np.asarray(S.data).astype(np.int64)), ctypes.c_int)
        a.tspan = ptr(da(m.tspan), ctypes.c_double)
        a.tlen = len(m.tspan)
        Nn, Nc, Nd, Nrep, tlen = m.Nn, m.Nc, m.Nd, m.replicates, len(m.tspan)
        self.U = self.V = self.prU = self.prV = None
        if m.U_sparse is None:
            self.U = np.zeros((Nn * Nc, Nrep * tlen), dtype=np.int32, order="F")
            a.U = ptr(self.U, ctypes.c_int)
        else:
            Us = m.U_sparse.tocsc()
            a.irU = ptr(ia(Us.indices), ctypes.c_int)
            a.jcU = ptr(ia(Us.indptr), ctypes.c_int)
            self.prU = np.zeros(Us.nnz, dtype=np.float64)
            a.prU = ptr(self.prU, ctypes.c_double)
        if m.V_sparse is None:
            self.V = np.zeros((Nn * Nd, Nrep * tlen), dtype=np.float64, order="F")
            a.V = ptr(self.V, ctypes.c_double)
        else:
            Vs = m.V_sparse.tocsc()
            a.irV = ptr(ia(Vs.indices), ctypes.c_int)
            a.jcV = ptr(ia(Vs.indptr), ctypes.c_int)
            self.prV = np.zeros(Vs.nnz, dtype=np.float64)
            a.prV = ptr(self.prV, ctypes.c_double)
        if m.Nld > 0:
            a.ldata = ptr(da(m.ldata), ctypes.c_double)
        else:
            # SimInf.c:238-244: ldata points at {INFINITY} when Nld == 0.
            a.ldata = ptr(da([np.inf]), ctypes.c_double)
        a.gdata = ptr(da(m.gdata if len(m.gdata) else [0.0]), ctypes.c_double)
        a.Nn, a.Nc, a.Nt, a.Nd, a.Nld, a.Nrep = Nn, Nc, m.Nt, Nd, m.Nld, Nrep
        a.irE = ptr(ia(E.indices), ctypes.c_int)
        a.jcE = ptr(ia(E.indptr), ctypes.c_int)
        a.prE = ptr(da(E.data), ctypes.c_double)
        if m.N is not None and m.N.shape[0] == Nc and m.N.size > 0:
            a.N = ptr(ia(m.N), ctypes.c_int)
        ev = m.events
        a.len = len(ev["event"])
        for key in ("event", "time", "node", "dest", "n", "select", "shift"):
            setattr(a, key, ptr(ia(ev[key]), ctypes.c_int))
        a.proportion = ptr(da(ev["proportion"]), ctypes.c_double)
        a.Nthread = 1
        a.solver = solver
        a.crn = crn
        a.rng_type = 0
        a.seed = seed
        self.args = a

"""ctypes bridge to the CPU checker.  TEST INFRASTRUCTURE ONLY: imported
by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import ctypes
import os
import time

import numpy as np

from siminf_amd import _native

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsiminf_oracle.so")

_lib = None
_host_plugins = {}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build
            build.build()
        lib = ctypes.CDLL(LIB)
        ap = ctypes.POINTER(_native.SolverArgs)
        for name in ("oracle_run", "oracle_run_solver_ssm", "oracle_run_solver_mssm",
                     "oracle_run_solver_mssm_crn", "oracle_run_solver_aem"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            f.argtypes = [ap]
        lib.oracle_set_model.restype = ctypes.c_int
        lib.oracle_set_model.argtypes = [ap, ctypes.c_char_p]
        lib.oracle_last_transitions.restype = ctypes.c_longlong
        lib.oracle_aem_ties.restype = ctypes.c_longlong
        u32p = ctypes.POINTER(ctypes.c_uint32)
        lib.oracle_philox.argtypes = [u32p, u32p, u32p]
        lib.oracle_uniforms.argtypes = [ctypes.c_ulonglong, ctypes.c_uint,
                                        ctypes.c_ulonglong, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_log.argtypes = [dp, dp, ctypes.c_longlong]
        lib.oracle_sqrt.argtypes = [dp, dp, ctypes.c_longlong]
        ip = ctypes.POINTER(ctypes.c_int)
        lib.oracle_binomial.argtypes = [ctypes.c_ulonglong, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ip]
        lib.oracle_hypergeometric.argtypes = [ctypes.c_ulonglong, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ip]
        lib.oracle_prefetch_mismatches.restype = ctypes.c_int
        lib.oracle_prefetch_mismatches.argtypes = [ctypes.c_ulonglong, ctypes.c_int,
                                                   ctypes.c_int]
        lib.oracle_pfilter.restype = ctypes.c_int
        lib.oracle_pfilter.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_int, ip, ctypes.c_ulonglong, dp, dp,
            ip]
        lib.oracle_exp.argtypes = [dp, dp, ctypes.c_longlong]
        lib.oracle_lfact.argtypes = [ip, dp, ctypes.c_longlong]
        lib.oracle_log_dbinom.argtypes = [ip, ip, ctypes.c_double, dp, ctypes.c_longlong]
        lib.oracle_log_dpois.argtypes = [ip, dp, dp, ctypes.c_longlong]
        lib.oracle_abc_weights.restype = ctypes.c_int
        lib.oracle_abc_weights.argtypes = [ctypes.c_char_p, dp, dp, ctypes.c_int,
                                           ctypes.c_int, dp, dp, dp, dp, dp]
        lib.oracle_set_replicate_data.restype = None
        lib.oracle_set_replicate_data.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int]
        lib.oracle_set_threads.restype = None
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_resample.restype = ctypes.c_int
        lib.oracle_sample.restype = ctypes.c_int
        lib.oracle_sample.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int, ip,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_ulonglong, ctypes.c_int, ip]
        _lib = lib
    return _lib


class OracleResult:
    def __init__(self, err, bundle, transitions, ties):
        self.err = err
        self.U = bundle.U
        self.V = bundle.V
        self.prU = bundle.prU
        self.prV = bundle.prV
        self.transitions = transitions
        self.aem_ties = ties


def _set_plugin(lib, model, b):
    """tr_fun / pts_fun of the model's host plugin into b.args."""
    if isinstance(model.plugin, str):
        if lib.oracle_set_model(ctypes.byref(b.args), model.plugin.encode()):
            raise ValueError("no host plugin for %s" % model.plugin)
    else:
        # A parsed model: the same functions compiled for the host.
        from siminf_amd import plugins
        so = plugins.compile_host(model.plugin.parsed)
        if so not in _host_plugins:
            _host_plugins[so] = ctypes.CDLL(so)
        h = _host_plugins[so]
        b.args.tr_fun = ctypes.cast(h.siminf_tr_fun, ctypes.c_void_p)
        b.args.pts_fun = ctypes.c_void_p.in_dll(h, "siminf_pts_fun")


def run(model, solver="ssm", seed=0, crn=False, threads=1, rep_ldata=None,
        rep_gdata=None):
    """Run the checker on the same SimInf_solver_args the GPU path gets;
    threads > 1 runs the node loops over that many OpenMP threads (node
    blocks, the reference's partition; the same results).  rep_ldata
    (Nld x Nn*Nrep) and rep_gdata (Ngdata x Nrep) give every replicate its
    own parameters, as Plan.set_replicate_data does."""
    lib = load()
    b = _native.ArgsBundle(model, solver=0 if solver == "ssm" else 1,
                           crn=1 if crn else 0, seed=int(seed))
    _set_plugin(lib, model, b)
    dp = ctypes.POINTER(ctypes.c_double)
    keep = []

    def arr(x, rows):
        if x is None:
            return None, 0
        a = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(rows, -1, order="F"))
        keep.append(a)
        return a.ctypes.data_as(dp), rows
    lp, _ = arr(rep_ldata, model.Nld)
    lnodes = 0
    if rep_ldata is not None:
        lnodes = model.Nn if np.size(rep_ldata) == model.Nld * model.Nn * model.replicates else 1
    gp, gs = arr(rep_gdata, len(model.gdata))
    lib.oracle_set_replicate_data(lp, lnodes, gp, gs)
    lib.oracle_set_threads(int(threads))
    try:
        t0 = time.perf_counter()
        err = lib.oracle_run(ctypes.byref(b.args))
        seconds = time.perf_counter() - t0
    finally:
        lib.oracle_set_threads(1)
        lib.oracle_set_replicate_data(None, 0, None, 0)
    r = OracleResult(err, b, lib.oracle_last_transitions(), lib.oracle_aem_ties())
    r.seconds = seconds      # the solver call alone
    return r


def pfilter(model, obs_compartment, times, obs, n_particles, obs_prob=0.5, seed=0,
            obs_dist="binomial", obs_offset=0.0):
    """The checker's filter on the same inputs as siminf_amd.pfilter;
    returns (error code, loglik, ess, ancestors)."""
    from siminf_amd import particle_filter as pf
    lib = load()
    dist = pf.OBS_DISTRIBUTIONS[obs_dist]
    tspan2, obs, obs_nodes = pf.prepare(model, obs_compartment, times, obs)
    b = _native.ArgsBundle(model, solver=0, seed=int(seed))
    _set_plugin(lib, model, b)
    n_int = len(tspan2)
    loglik = ctypes.c_double()
    ess = np.zeros(n_int)
    anc = np.zeros((n_int, int(n_particles)), dtype=np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    err = lib.oracle_pfilter(ctypes.byref(b.args), int(n_particles), n_int,
                             tspan2.ctypes.data_as(dp), int(obs_compartment), dist,
                             float(obs_prob), float(obs_offset), obs_nodes,
                             obs.ctypes.data_as(ip),
                             ctypes.c_ulonglong(int(seed)), ctypes.byref(loglik),
                             ess.ctypes.data_as(dp), anc.ctypes.data_as(ip))
    return err, loglik.value, ess, anc


def _arr(fn, x, *extra, dtype=np.float64, out_dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=dtype)
    y = np.empty(x.shape, dtype=out_dtype)
    ptype = ctypes.POINTER(ctypes.c_double if dtype == np.float64 else ctypes.c_int)
    fn(x.ctypes.data_as(ptype), *extra, y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
       x.size)
    return y


def exp(x):
    return _arr(load().oracle_exp, x)


def lfact(n):
    return _arr(load().oracle_lfact, n, dtype=np.int32)


def log_dbinom(x, n, p):
    x = np.ascontiguousarray(x, dtype=np.int32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    y = np.empty(x.shape)
    ip = ctypes.POINTER(ctypes.c_int)
    load().oracle_log_dbinom(x.ctypes.data_as(ip), n.ctypes.data_as(ip), ctypes.c_double(p),
                             y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), x.size)
    return y


def log_dpois(x, lam):
    x = np.ascontiguousarray(x, dtype=np.int32)
    lam = np.ascontiguousarray(np.broadcast_to(lam, x.shape), dtype=np.float64)
    y = np.empty(x.shape)
    dp = ctypes.POINTER(ctypes.c_double)
    load().oracle_log_dpois(x.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            lam.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size)
    return y


def abc_weights(distribution, p1, p2, x, xx, w, sigma):
    """The checker's ABC weights; x, xx: (particles x parameters);
    returns (error code, weights)."""
    from siminf_amd import abc_smc as A
    return A.call_weights(load().oracle_abc_weights, distribution, p1, p2, x, xx, w, sigma)


def resample(w, u01):
    w = np.ascontiguousarray(w, dtype=np.float64)
    idx = np.zeros(len(w), dtype=np.int32)
    err = load().oracle_resample(w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 len(w), ctypes.c_double(u01),
                                 idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return err, idx


def uniforms(seed, domain, sid, count):
    out = np.zeros(count, dtype=np.float64)
    load().oracle_uniforms(seed, domain, sid, count,
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    load().oracle_philox(c, k, o)
    return list(o)


def log(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    load().oracle_log(x.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size)
    return y


def sqrt(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    load().oracle_sqrt(x.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size)
    return y


def binomial(seed, p, n, count):
    out = np.zeros(count, dtype=np.int32)
    load().oracle_binomial(seed, p, n, count,
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out


def hypergeometric(seed, n1, n2, t, count):
    out = np.zeros(count, dtype=np.int32)
    load().oracle_hypergeometric(seed, n1, n2, t, count,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out


def prefetch_mismatches(seed, npre, count):
    return load().oracle_prefetch_mismatches(seed, npre, count)


def sample(E, u_node, select, n, proportion, seed, count, enter=False):
    """Repeated sample_select / sample_select_enter on one node; returns
    (error code, count x Nc array).  select is zero-based."""
    import scipy.sparse as sp
    E = sp.csc_matrix(E)
    E.sort_indices()
    ir = np.ascontiguousarray(E.indices, dtype=np.int32)
    jc = np.ascontiguousarray(E.indptr, dtype=np.int32)
    pr = np.ascontiguousarray(E.data, dtype=np.float64)
    u = np.ascontiguousarray(u_node, dtype=np.int32)
    Nc = E.shape[0]
    out = np.zeros((count, Nc), dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    err = load().oracle_sample(int(enter), ir.ctypes.data_as(ip), jc.ctypes.data_as(ip),
                               pr.ctypes.data_as(dp), Nc, u.ctypes.data_as(ip),
                               select, n, proportion, seed, count,
                               out.ctypes.data_as(ip))
    return err, out

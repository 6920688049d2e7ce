"""A model resident on the GPU: inputs uploaded once, runs with new seeds,
and summaries of the trajectory computed where it lies.

`Plan` wraps the `SimInf_hip_plan_*` entry points (no reference
counterpart; they serve callers that run one model many times).  Its
`prevalence` and `trajectory` mirror the reference's prevalence() and
trajectory(format = "matrix") (R/prevalence.R:70-110,
SimInf_trajectory.c:566-650) on the device-resident U, so a summary costs
a kernel and a copy of the summary, never a copy of U.
"""
import ctypes

import numpy as np

from . import _native, summaries
from .model import SimInfError, _attach, _raise, device_plugin


class Plan:
    def __init__(self, model, solver="ssm", lib=None):
        self.lib = lib or _native.load()
        if solver not in ("ssm", "aem"):
            _raise(self.lib, -12)
        self.model = model
        self._bundle = _native.ArgsBundle(model, solver=0 if solver == "ssm" else 1, seed=1)
        self._handle = ctypes.c_void_p()
        err = self.lib.SimInf_hip_plan_create(ctypes.byref(self._handle),
                                              ctypes.byref(self._bundle.args),
                                              ctypes.c_void_p(device_plugin(model, self.lib)))
        if err:
            _raise(self.lib, err)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self._handle:
            self.lib.SimInf_hip_plan_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, err):
        if err:
            try:
                _raise(self.lib, err)
            except SimInfError as e:
                e.detail = self.last_error()
                raise

    def last_error(self):
        """Where the last run failed: {"replicate", "day", "phase" and
        "node" or "event" (zero-based, the event's index in the model's
        sorted event list)}, or None if it succeeded."""
        c, r, d, ph = (ctypes.c_int() for _ in range(4))
        w = ctypes.c_longlong()
        if self.lib.SimInf_hip_plan_last_error(self._handle, ctypes.byref(c), ctypes.byref(r),
                                               ctypes.byref(d), ctypes.byref(ph),
                                               ctypes.byref(w)) or c.value == 0:
            return None
        phase = SimInfError.PHASES[ph.value]
        where = "event" if ph.value in (2, 3) else "node"
        return {"code": c.value, "replicate": r.value, "day": d.value, "phase": phase,
                where: w.value}

    def run(self, seed):
        """One run with this seed; returns (transitions, node_days,
        kernel_ms)."""
        self._check(self.lib.SimInf_hip_plan_run(self._handle,
                                                 int(seed) & 0xFFFFFFFFFFFFFFFF))
        tr, nd, ms = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        self.lib.SimInf_hip_plan_stats(self._handle, ctypes.byref(tr), ctypes.byref(nd),
                                       ctypes.byref(ms))
        return tr.value, nd.value, ms.value

    def set_replicate_data(self, ldata=None, gdata=None):
        """Parameters per replicate: ldata (Nld x Nn*Nrep, column
        replicate*Nn + node, or Nld x Nrep for the same values on every
        node of a replicate) and/or gdata (Ngdata x Nrep); later runs read
        them."""
        dp = ctypes.POINTER(ctypes.c_double)
        keep = []

        def arr(x, shape):
            if x is None:
                return None
            a = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(shape, order="F"))
            keep.append(a)
            return a.ctypes.data_as(dp)
        m = self.model
        nodes = 0
        if ldata is not None:
            nodes = m.Nn if np.size(ldata) == m.Nld * m.Nn * m.replicates else 1
        lp = arr(ldata, (m.Nld, nodes * m.replicates))
        gp = arr(gdata, (len(m.gdata), m.replicates))
        self._check(self.lib.SimInf_hip_plan_set_replicate_data(self._handle, lp, nodes, gp))

    def set_lanes_per_wave(self, shift):
        """2^shift lanes per wavefront in the node kernels' active pass (-1:
        the default rule)."""
        self._check(self.lib.SimInf_hip_plan_set_lanes_per_wave(self._handle, int(shift)))

    def set_fill_block(self, cols):
        """Columns each fill-kernel workgroup writes per 256 nodes (the
        trajectory-flush schedule; 0: the default)."""
        self._check(self.lib.SimInf_hip_plan_set_fill_block(self._handle, int(cols)))

    def fill_block(self):
        """Columns per fill block in use."""
        c = ctypes.c_int()
        self._check(self.lib.SimInf_hip_plan_get_fill_block(self._handle, ctypes.byref(c)))
        return c.value

    def set_herd_map(self, solo=(), herds_per_wavefront=64):
        """Grid-wide coupled plans: each herd in solo (zero-based, at most
        16) on a wavefront of its own, the others herds_per_wavefront (8 to
        64) to a wavefront."""
        nodes = np.ascontiguousarray(np.asarray(solo, dtype=np.int32).ravel())
        self._check(self.lib.SimInf_hip_plan_set_herd_map(
            self._handle, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(nodes),
            int(herds_per_wavefront)))

    def herd_map(self):
        """(herds on their own wavefront, herds per packed wavefront)."""
        n, k = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.SimInf_hip_plan_get_herd_map(self._handle, ctypes.byref(n),
                                                          ctypes.byref(k)))
        return n.value, k.value

    def autotune(self, seed=1, reps=2):
        """Try every candidate launch shape, keep the fastest; returns
        (best, mean solver ms per candidate): the lanes-per-wavefront shift
        0..6 for node kernels (7 times); for grid-wide coupled kernels the
        number of busiest herds kept on wavefronts of their own (8 times:
        4 solo herds with the others 64, 32, 16, 8 to a wavefront, then 0,
        2, 8, 16 solo herds with the best packing; herd_map() gives the
        kept map); -1 when there is nothing to tune."""
        best = ctypes.c_int()
        ms = (ctypes.c_double * 8)(*([float("nan")] * 8))
        self._check(self.lib.SimInf_hip_plan_autotune(self._handle, int(seed), int(reps),
                                                      ctypes.byref(best), ms))
        times = list(ms)
        while times and times[-1] != times[-1]:     # entries the plan did not use
            times.pop()
        return best.value, times

    def download(self):
        """Copy the last run's U/V into the model (as run() leaves them)."""
        b = self._bundle
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        self._check(self.lib.SimInf_hip_plan_download(
            self._handle,
            None if b.U is None else _native.ptr(b.U, ctypes.c_int),
            None if b.V is None else _native.ptr(b.V, ctypes.c_double),
            None if b.prU is None else b.prU.ctypes.data_as(dp),
            None if b.prV is None else b.prV.ctypes.data_as(dp)))
        _attach(self.model, b)
        return self.model

    def _columns(self):
        return self.model.replicates * len(self.model.tspan)

    def fuse_prevalence(self, formula):
        """Keep, from the next run on, only what prevalence(formula,
        level=1) needs (exact sums per replicate and column) instead of U
        and V; returns False (and changes nothing) for plans that cannot
        fuse.  A fused plan's download and trajectories raise."""
        cases, pop, ops, _, _ = summaries.prevalence_spec(formula, self.model.compartments, 1)
        if len(ops):
            _raise(self.lib, -105)
        err = self.lib.SimInf_hip_plan_fuse_prevalence(self._handle, cases, pop)
        if err == 1:
            return False
        self._check(err)
        return True

    def prevalence(self, formula, level=1, index=None):
        """prevalence(model, formula, level, index, format = "matrix") of the
        last run, computed on the device.  index is one-based like the
        reference's.  Returns Nrep*tlen values (levels 1, 2) or a
        len(index) x Nrep*tlen matrix (level 3)."""
        cases, pop, ops, args, level = summaries.prevalence_spec(
            formula, self.model.compartments, level)
        idx = summaries.node_index(index, self.model.Nn)
        n_sel = self.model.Nn if idx is None else len(idx)
        ncols = self._columns()
        out = np.zeros((n_sel, ncols) if level == 3 else ncols, dtype=np.float64, order="F")
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        self._check(self.lib.SimInf_hip_plan_prevalence(
            self._handle, cases, pop, ops.ctypes.data_as(ip), args.ctypes.data_as(dp),
            len(ops), level, None if idx is None else idx.ctypes.data_as(ip),
            0 if idx is None else len(idx), out.ctypes.data_as(dp)))
        return out

    def trajectory(self, compartments=None, index=None):
        """trajectory(model, compartments, index, format = "matrix") of the
        last run, gathered on the device: rows node-major then compartment
        (model order), one column per U column.  Selecting only continuous
        states (names of v0's rows) gathers them from V instead, as float64
        (R/trajectory.R:220-228); U and V names together raise the
        reference's message."""
        names = summaries.selected_names(compartments)
        vnames = [n for n in (names or []) if n in self.model.v0_names
                  and n not in self.model.compartments]
        if vnames:
            bad = [n for n in names if n not in self.model.compartments
                   and n not in self.model.v0_names]
            if bad:
                raise summaries.SummaryError("Non-existing compartment(s) in model: %s." %
                                             ", ".join("'%s'" % x for x in bad))
            if len(vnames) != len(names):
                raise summaries.SummaryError("Cannot combine data from different slots.")
            return self._trajectory_v(vnames, index)
        m = summaries.trajectory_mask(compartments, self.model.compartments)
        idx = summaries.node_index(index, self.model.Nn)
        n_sel = self.model.Nn if idx is None else len(idx)
        ncomp = bin(m).count("1")
        out = np.zeros((n_sel * ncomp, self._columns()), dtype=np.int32, order="F")
        ip = ctypes.POINTER(ctypes.c_int)
        self._check(self.lib.SimInf_hip_plan_trajectory(
            self._handle, m, None if idx is None else idx.ctypes.data_as(ip),
            0 if idx is None else len(idx), out.ctypes.data_as(ip)))
        return out

    def _trajectory_v(self, names, index):
        m = summaries.mask(names, self.model.v0_names)
        idx = summaries.node_index(index, self.model.Nn)
        n_sel = self.model.Nn if idx is None else len(idx)
        out = np.zeros((n_sel * len(names), self._columns()), dtype=np.float64, order="F")
        self._check(self.lib.SimInf_hip_plan_trajectory_v(
            self._handle, m, None if idx is None else idx.ctypes.data_as(
                ctypes.POINTER(ctypes.c_int)),
            0 if idx is None else len(idx), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out


class Batch:
    """Several plans -- any models, solvers and sizes -- run as one graph
    launch (SimInf_hip_batch_*; no reference counterpart).  Each plan's
    results after run() are those of its own Plan.run with the same seed.
    The plans must outlive the batch."""

    def __init__(self, plans, lib=None):
        self.lib = lib or _native.load()
        self.plans = list(plans)
        handles = (ctypes.c_void_p * len(self.plans))(*[p._handle for p in self.plans])
        self._handle = ctypes.c_void_p()
        err = self.lib.SimInf_hip_batch_create(ctypes.byref(self._handle), handles,
                                               len(self.plans))
        if err:
            _raise(self.lib, err)

    def set_graph(self, use_graph):
        """True: each run is one graph launch (the default); False: the
        same kernels enqueued directly on the plans' streams."""
        err = self.lib.SimInf_hip_batch_set_graph(self._handle, int(bool(use_graph)))
        if err:
            _raise(self.lib, err)

    def autotune(self, seed=1, reps=2):
        """Tune each plan's lanes per wavefront and columns per fill block
        in the batch, and the graph against the direct enqueue; returns
        (mean launch ms of the kept shape, [shift per plan], uses_graph);
        Plan.fill_block() gives each plan's kept fill block."""
        ms = ctypes.c_double()
        err = self.lib.SimInf_hip_batch_autotune(self._handle, int(seed), int(reps),
                                                 ctypes.byref(ms))
        if err:
            _raise(self.lib, err)
        shifts, g = [], ctypes.c_int()
        for i in range(len(self.plans)):
            sh = ctypes.c_int()
            self.lib.SimInf_hip_batch_get_shape(self._handle, i, ctypes.byref(sh),
                                                ctypes.byref(g))
            shifts.append(sh.value)
        return ms.value, shifts, bool(g.value)

    def run(self, seeds):
        """One launch, plan i with seeds[i]; returns the launch's device
        time in milliseconds."""
        sd = (ctypes.c_ulong * len(self.plans))(
            *[int(x) & 0xFFFFFFFFFFFFFFFF for x in seeds])
        err = self.lib.SimInf_hip_batch_run(self._handle, sd)
        if err:
            _raise(self.lib, err)
        ms = ctypes.c_double()
        self.lib.SimInf_hip_batch_time(self._handle, ctypes.byref(ms))
        return ms.value

    def close(self):
        if self._handle:
            self.lib.SimInf_hip_batch_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Plan", "Batch", "SimInfError"]

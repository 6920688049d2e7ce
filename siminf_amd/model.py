"""Host-side model layer: the layouts SimInf's R layer hands to the C
boundary, and run() over the HIP solver core.

Mirrors (what, not how):
  SimInf_model slots and validity   R/SimInf_model.R:158-206, R/valid.R
  SimInf_events (sorted events)     R/SimInf_events.R:367-413
  run(model, solver, seed, crn)     R/run.R:190-283, src/SimInf.c:136-409
  built-in constructors             R/SIR.R:110-206, R/SEIR.R:117-158,
                                    R/SIS.R, R/SISe.R:99-199,
                                    R/SISe3.R:112-205

Matrices use the reference's memory layouts: u0 is int Nc x (Nn*Nrep)
(one column per node), v0 double Nd x (Nn*Nrep), ldata Nld x Nn, and
U/V are (Nn*Nc) x (Nrep*tlen) column-major.  S, G and E are
scipy.sparse CSC matrices (the dgCMatrix slots i/p/x); N is a dense int
matrix Nc x Nshift.
"""
import numpy as np
import scipy.sparse as sp

from . import _native

EVENT_TYPES = {"exit": 0, "enter": 1, "intTrans": 2, "extTrans": 3}


class SimInfError(RuntimeError):
    """An error code of the solver core, with the reference's message.
    Errors of a plan's run also carry where it failed (``detail``: a dict
    of replicate, day, phase and node or event), or None."""

    PHASES = ("initial rates", "ssa", "scheduled events", "external transfers",
              "post time step")

    def __init__(self, code, message, detail=None):
        super().__init__(message)
        self.code = code
        self.detail = detail


def _csc(x, dtype=np.float64):
    if sp.issparse(x):
        m = sp.csc_matrix(x, dtype=dtype)
    else:
        m = sp.csc_matrix(np.asarray(x, dtype=dtype))
    m.sort_indices()
    return m


def make_events(E, N=None, events=None):
    """Event list sorted by (time, event, select), stably, as
    SimInf_events does (R/SimInf_events.R:398).  ``events`` maps the
    column names event/time/node/dest/n/proportion/select/shift to
    arrays (event may use the names exit/enter/intTrans/extTrans)."""
    cols = ("event", "time", "node", "dest", "n", "proportion", "select", "shift")
    if events is None:
        ev = {c: np.zeros(0, dtype=np.float64 if c == "proportion" else np.int32)
              for c in cols}
    else:
        ev = {}
        n = len(np.asarray(events["event"]))
        for c in cols:
            x = events.get(c, np.zeros(n)) if isinstance(events, dict) else events[c]
            x = np.asarray(x)
            if c == "event" and x.dtype.kind in "US":
                x = np.array([EVENT_TYPES[str(v)] for v in x])
            if c == "proportion":
                ev[c] = x.astype(np.float64)
            else:
                ev[c] = x.astype(np.int64).astype(np.int32)
        order = np.lexsort((ev["select"], ev["event"], ev["time"]))
        ev = {c: np.ascontiguousarray(v[order]) for c, v in ev.items()}
    ev["E"] = _csc(E if E is not None else np.zeros((0, 0)))
    ev["N"] = None if N is None else np.asarray(N, dtype=np.int32)
    return ev


class SimInfModel:
    """A compartment model in the layouts of the C boundary.

    ``plugin`` names the device plugin (``SimInf_hip_model_builtin``) that
    holds the model's transition rate and post-time-step functions.
    """

    def __init__(self, plugin, G, S, tspan, u0, ldata=None, gdata=None,
                 v0=None, E=None, N=None, events=None, replicates=1,
                 U_sparse=None, V_sparse=None, compartments=None, v0_names=None,
                 ldata_names=None, gdata_names=None):
        self.plugin = plugin
        self.G = _csc(G)
        self.S = _csc(S)
        self.tspan = np.asarray(tspan, dtype=np.float64).ravel()
        self.replicates = int(replicates)
        u0 = np.asarray(u0)
        if u0.ndim == 1:
            u0 = u0.reshape(-1, 1)
        self.Nc = self.S.shape[0]
        self.Nt = self.S.shape[1]
        if u0.shape[0] != self.Nc:
            raise ValueError("u0 must have one row per compartment")
        # The rownames of S: the compartment names.
        if compartments is None:
            compartments = ["C%d" % (c + 1) for c in range(self.Nc)]
        if len(compartments) != self.Nc:
            raise ValueError("one name per compartment")
        self.compartments = list(compartments)
        self._u0 = np.asfortranarray(u0.astype(np.int32))
        ncols = self._u0.shape[1]
        if ncols % self.replicates:
            raise ValueError("ncol(u0) must be a multiple of replicates")
        self.Nn = ncols // self.replicates
        if v0 is None:
            v0 = np.zeros((0, ncols))
        v0 = np.asarray(v0, dtype=np.float64)
        if v0.ndim == 1:
            v0 = v0.reshape(1, -1)
        self._v0 = np.asfortranarray(v0)
        self.Nd = self._v0.shape[0]
        # The rownames of v0.
        if v0_names is None:
            v0_names = ["V%d" % (d + 1) for d in range(self.Nd)]
        if len(v0_names) != self.Nd:
            raise ValueError("one name per continuous state")
        self.v0_names = list(v0_names)
        if ldata is None:
            ldata = np.zeros((0, self.Nn))
        ldata = np.asarray(ldata, dtype=np.float64)
        if ldata.ndim == 1:
            ldata = ldata.reshape(-1, 1)
        self.ldata = np.asfortranarray(ldata)
        self.Nld = self.ldata.shape[0]
        self.gdata = np.asarray(gdata if gdata is not None else [],
                                dtype=np.float64).ravel()
        # Row names of ldata and names of gdata (the parameters abc() can
        # infer).
        self.ldata_names = list(ldata_names) if ldata_names is not None else \
            ["L%d" % (k + 1) for k in range(self.Nld)]
        self.gdata_names = list(gdata_names) if gdata_names is not None else \
            ["G%d" % (k + 1) for k in range(len(self.gdata))]
        if len(self.ldata_names) != self.Nld or len(self.gdata_names) != len(self.gdata):
            raise ValueError("one name per ldata row and per gdata value")
        self.events = make_events(E if E is not None else np.zeros((self.Nc, 0)),
                                  N, events)
        self.E = self.events["E"]
        self.N = self.events["N"]
        self.U_sparse = U_sparse
        self.V_sparse = V_sparse
        self.U = None
        self.V = None

    def u0_full(self):
        return self._u0

    def v0_full(self):
        return self._v0

    @property
    def u0(self):
        return self._u0

    @property
    def v0(self):
        return self._v0


def run(model, solver="ssm", seed=None, crn=False, lib=None):
    """Run the model on the GPU; mirrors run() (R/run.R:190-283) and
    SimInf_run's option parsing and dispatch (src/SimInf.c:136-409).

    Returns the model with ``U``/``V`` (dense) or ``U_sparse``/``V_sparse``
    values filled.  Errors raise SimInfError with the reference's message.
    """
    lib = lib or _native.load()
    if solver not in ("ssm", "aem"):
        _raise(lib, -12)
    if seed is None:
        seed = int(np.random.randint(0, 2**32 - 1))
    if isinstance(seed, (bool, np.bool_)) or not np.isfinite(float(seed)):
        _raise(lib, -22)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    dev = device_plugin(model, lib)
    b = _native.ArgsBundle(model, solver=0 if solver == "ssm" else 1,
                           crn=1 if crn else 0, seed=seed)
    import ctypes
    err = lib.SimInf_hip_run(ctypes.byref(b.args), ctypes.c_void_p(dev))
    if err:
        _raise(lib, err)
    _attach(model, b)
    return model


def device_plugin(model, lib=None):
    """SimInf_hip_model* for the model: a built-in plugin by name, or the
    compiled plugin of a parsed model."""
    if isinstance(model.plugin, str):
        lib = lib or _native.load()
        dev = lib.SimInf_hip_model_builtin(model.plugin.encode())
        if not dev:
            raise SimInfError(-9, "Unknown device plugin '%s'." % model.plugin)
        return dev
    return model.plugin.device_model()


def _attach(model, b):
    if b.U is not None:
        model.U = b.U
    else:
        model.U_sparse = model.U_sparse.tocsc().copy()
        model.U_sparse.data = b.prU.copy()
    if b.V is not None:
        model.V = b.V
    else:
        model.V_sparse = model.V_sparse.tocsc().copy()
        model.V_sparse.data = b.prV.copy()


def _raise(lib, code):
    raise SimInfError(code, lib.SimInf_hip_error_message(code).decode())


def trajectory(model, compartment_index, node_index=None):
    """Dense U slice: rows of one compartment over time, shape
    (len(nodes), Nrep*tlen); compartments/nodes are zero-based."""
    U = model.U
    Nc = model.Nc
    nodes = np.arange(model.Nn) if node_index is None else np.asarray(node_index)
    return U[nodes * Nc + compartment_index, :]

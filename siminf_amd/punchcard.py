"""Sparse output templates: `punchcard(model) <- value`
(R/punchcard.R:128-256).

``value`` is a data frame (a pandas DataFrame or a dict of equal-length
columns) with the columns ``time`` and ``node`` (one-based) and optionally
one logical column per compartment or continuous state to record; a
missing or NA entry is FALSE, and with neither kind of column every
compartment and state is recorded.  The result is the pattern of
`U_sparse` / `V_sparse` that the solver's sparse writeback fills
(SimInf_solver.c:930-950, replicate r's columns offset by r * tlen), or a
dense output where every node, compartment and time point is selected.
"""
import numpy as np
import scipy.sparse as sp


class PunchcardError(ValueError):
    """A template the reference rejects (same message)."""


def _columns(value):
    try:
        import pandas as pd
        if isinstance(value, pd.DataFrame):
            return {c: value[c].to_numpy() for c in value.columns}
    except ImportError:
        pass
    if isinstance(value, dict):
        cols = {k: np.atleast_1d(np.asarray(v)) for k, v in value.items()}
        n = max((len(v) for v in cols.values()), default=0)
        return {k: (np.resize(v, n) if len(v) != n else v) for k, v in cols.items()}
    raise PunchcardError("'value' argument is not a 'data.frame'.")


def _logical(x):
    """as.logical with NA -> FALSE."""
    x = np.asarray(x)
    if x.dtype == object:
        return np.array([bool(v) if v is not None and v == v else False for v in x])
    if x.dtype.kind == "f":
        return np.where(np.isnan(x), False, x != 0)
    return x.astype(bool)


def create_template(value, tspan, n_nodes, names, replicates):
    """(sparse pattern or None for dense output) of one slot
    (create_template, R/punchcard.R:169-256).  names: the slot's row names
    (compartments or continuous states)."""
    dims = (n_nodes * len(names), replicates * len(tspan))
    if value is None:
        return None
    cols = _columns(value)
    nrow = max((len(v) for v in cols.values()), default=0)
    empty = sp.csc_matrix(dims, dtype=np.float64)
    if nrow == 0:
        return empty
    if "node" not in cols or "time" not in cols:
        raise PunchcardError("'value' must have the columns 'time' and 'node'.")
    time = np.asarray(cols["time"], dtype=np.float64)
    node = np.asarray(cols["node"], dtype=np.float64)
    order = np.lexsort((node, time))
    time, node = time[order], node[order]
    if not np.all((node == np.round(node)) & (node >= 1) & (node <= n_nodes)):
        raise PunchcardError("Unable to match all nodes.")
    tpos = {t: k for k, t in reversed(list(enumerate(np.asarray(tspan, dtype=float))))}
    if not all(t in tpos for t in time):
        raise PunchcardError("Unable to match all time-points to tspan.")
    i = node.astype(np.int64) - 1
    j = np.array([tpos[t] for t in time], dtype=np.int64)
    selected = [c for c in cols if c not in ("time", "node")]
    if not selected:
        flags = np.ones((nrow, len(names)), dtype=bool)
    elif any(c in names for c in selected):
        flags = np.zeros((nrow, len(names)), dtype=bool)
        for k, c in enumerate(names):
            if c in cols:
                flags[:, k] = _logical(cols[c])[order]
    else:
        return empty
    rows = (i[:, None] * len(names) + np.arange(len(names))[None, :])[flags]
    colj = np.repeat(j, len(names)).reshape(nrow, len(names))[flags]
    if flags.sum() == n_nodes * len(names) * len(tspan):
        return None
    rows = np.tile(rows, replicates)
    colj = np.concatenate([colj + r * len(tspan) for r in range(replicates)])
    m = sp.coo_matrix((np.ones(len(rows)), (rows, colj)), shape=dims).tocsc()
    m.sum_duplicates()
    m.sort_indices()
    m.data[:] = np.nan
    return m


def punchcard(model, value):
    """Set the model's sparse output templates (punchcard<-,
    R/punchcard.R:128-153); value None restores dense output."""
    model.U_sparse = create_template(value, model.tspan, model.Nn, model.compartments,
                                     model.replicates)
    model.V_sparse = create_template(value, model.tspan, model.Nn, model.v0_names,
                                     model.replicates)
    model.U = model.V = None
    return model

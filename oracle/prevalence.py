"""Numpy restatement of the reference's prevalence() and
trajectory(format = "matrix").  TEST INFRASTRUCTURE ONLY: imported by
tests/ as the checker of the device summaries.

  sum_compartments      R/prevalence.R:24-38
  evaluate_condition    R/prevalence.R:40-68 (the condition is a Python
                        function of a dict of compartment arrays, written
                        beside the R text it restates)
  calculate_prevalence  R/prevalence.R:70-110
  trajectory matrix     SimInf_trajectory.c:566-650 (rows node-major, then
                        compartment in model order)
"""
import numpy as np


def _block(U, Nc, compartment, index):
    """Rows of one compartment for the selected nodes: (nodes x columns)."""
    rows = U[compartment::Nc, :]
    return rows if index is None else rows[np.asarray(index), :]


def prevalence(U, compartments, cases, population, level=1, index=None,
               condition=None):
    """U: (Nn*Nc) x columns int matrix; cases / population: compartment
    names; index: zero-based nodes or None; condition: function of
    {name: nodes x columns int array} returning a bool array, or None."""
    Nc = len(compartments)
    pos = {c: k for k, c in enumerate(compartments)}
    x = sum(_block(U, Nc, pos[c], index).astype(np.int64) for c in cases)
    y = sum(_block(U, Nc, pos[c], index).astype(np.int64) for c in population)
    if condition is not None:
        env = {c: _block(U, Nc, pos[c], index).astype(np.int64) for c in compartments}
        k = np.asarray(condition(env), dtype=bool)
        assert k.shape == x.shape
        x = x * k
        y = y * k
    with np.errstate(invalid="ignore", divide="ignore"):
        if level == 1:
            return x.sum(axis=0).astype(np.float64) / y.sum(axis=0).astype(np.float64)
        if level == 2:
            return ((x > 0).sum(axis=0).astype(np.float64) /
                    (y > 0).sum(axis=0).astype(np.float64))
        return x.astype(np.float64) / y.astype(np.float64)


def trajectory(U, compartments, selected, index=None):
    Nc = len(compartments)
    Nn = U.shape[0] // Nc
    nodes = np.arange(Nn) if index is None else np.asarray(index)
    rows = [node * Nc + c for node in nodes for c in range(Nc)
            if compartments[c] in selected]
    return U[rows, :]

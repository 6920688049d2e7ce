"""Spatial coupling of the _sp models: the distance matrix between nodes
and the neighbour lists appended to ldata.

  distance_matrix                 R/distance.R:76-83,
                                  src/misc/SimInf_distance_matrix.c
  ldata_sp                        src/misc/SimInf_ldata.c (SimInf_ldata_sp)
  add_spatial_coupling_to_ldata   R/distance.R:156-176

A distance matrix is returned as a scipy CSC matrix with the reference's
dgCMatrix layout: column i lists node i's neighbours j != i within the
cutoff, in increasing j, with their Euclidean distance.  ldata_sp appends
to each node's ldata column its (neighbour index, value) pairs with
0-based indices, then the stop index -1, padded with zeros to the longest
list plus its stop.  R's NA has no counterpart in numpy; NaN coordinates
play its part here.
"""
import math

import numpy as np
import scipy.sparse as sp


class SpatialError(ValueError):
    """Raised with the reference's message for invalid spatial input."""


def distance_matrix(x, y, cutoff, min_dist=None, na_fail=True):
    """Distances between the points (x[i], y[i]) that are at most cutoff
    apart (SimInf_distance_matrix.c: SimInf_Euclidean_distance).  A zero
    distance becomes min_dist, which must then be given and positive."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    y = np.atleast_1d(np.asarray(y, dtype=np.float64))
    n = len(x)
    if n < 1:
        raise SpatialError("'x' must be a numeric vector with length > 0.")
    if len(y) != n:
        raise SpatialError("'y' must be a numeric vector with length %d." % n)
    cutoff = float(cutoff)
    if not math.isfinite(cutoff) or cutoff < 0:
        raise SpatialError("'cutoff' must be > 0.")
    if not isinstance(na_fail, (bool, np.bool_)):
        raise SpatialError("'na_fail' must be TRUE or FALSE.")
    try:
        md = float("nan") if min_dist is None else float(min_dist)
    except (TypeError, ValueError):
        md = float("nan")    # as.numeric of a non-number is NA
    if n > 2048 and np.isfinite(x).all() and np.isfinite(y).all():
        return _distance_matrix_tree(x, y, cutoff, md)
    rows, vals, ptr = [], [], [0]
    idx = np.arange(n)
    for i in range(n):
        d = np.hypot(x[i] - x, y[i] - y)
        d[i] = 0.0
        bad = ~np.isfinite(d)
        bad[i] = False
        if bad.any():
            na = np.isnan(x) | np.isnan(y)
            skip = bad & (na | na[i]) & (not na_fail)
            bad &= ~skip
            d[skip] = np.inf
        near = (d <= cutoff) & (idx != i)
        zero = near & (d <= 0)
        if not (math.isfinite(md) and md >= 0):
            # The first offending j in the reference's loop order decides.
            j_bad = int(np.flatnonzero(bad)[0]) if bad.any() else n
            j_zero = int(np.flatnonzero(zero)[0]) if zero.any() else n
            if j_zero < j_bad:
                raise SpatialError("Invalid 'min_dist' argument. "
                                   "Please provide 'min_dist' > 0.")
        if bad.any():
            j = int(np.flatnonzero(bad)[0])
            raise SpatialError("Invalid distance for i=%d and j=%d." % (i, j))
        dj = np.where(d[near] <= 0, md, d[near])
        rows.append(idx[near])
        vals.append(dj)
        ptr.append(ptr[-1] + len(dj))
    return sp.csc_matrix((np.concatenate(vals), np.concatenate(rows).astype(np.int32),
                          np.asarray(ptr, dtype=np.int32)), shape=(n, n))


def _distance_matrix_tree(x, y, cutoff, md):
    """distance_matrix for many finite points: a k-d tree proposes the
    pairs within a slightly larger radius, and each pair is then kept or
    dropped by the reference's own test on hypot (so pairs at the cutoff
    fall as the reference's do)."""
    from scipy.spatial import cKDTree
    n = len(x)
    tree = cKDTree(np.column_stack([x, y]))
    pairs = tree.query_pairs(cutoff * (1.0 + 1e-9) + 1e-300, output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1]])      # column (node)
    j = np.concatenate([pairs[:, 1], pairs[:, 0]])      # row (neighbour)
    d = np.hypot(x[i] - x[j], y[i] - y[j])
    keep = d <= cutoff
    i, j, d = i[keep], j[keep], d[keep]
    if (d <= 0).any():
        if not (math.isfinite(md) and md >= 0):
            raise SpatialError("Invalid 'min_dist' argument. "
                               "Please provide 'min_dist' > 0.")
        d = np.where(d <= 0, md, d)
    order = np.lexsort((j, i))
    i, j, d = i[order], j[order], d[order]
    ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(i, minlength=n), out=ptr[1:])
    return sp.csc_matrix((d, j.astype(np.int32), ptr), shape=(n, n))


def ldata_sp(data, distance, metric):
    """ldata of a spatial model: each node's column of data, then its
    neighbours' (index, value) pairs and the stop -1 (SimInf_ldata_sp).
    The value is the distance (metric 1), 1 / distance^2 (metric 2) or the
    neighbour's degree (any other metric)."""
    if data is None or isinstance(data, (list, tuple)):
        raise SpatialError("Invalid 'data' argument.")
    data = np.asarray(data, dtype=np.float64) if not sp.issparse(data) else None
    if data is None or data.ndim != 2:
        raise SpatialError("Invalid 'data' argument.")
    if not sp.issparse(distance) or distance.shape[0] != distance.shape[1]:
        raise SpatialError("Invalid 'distance' argument.")
    if isinstance(metric, (bool, np.bool_)) or not isinstance(metric, (int, np.integer)):
        raise SpatialError("Invalid 'metric' argument.")
    d = sp.csc_matrix(distance)
    nn = data.shape[1]
    if nn != d.shape[1]:
        raise SpatialError("The number of nodes in 'data' and 'distance' are not equal.")
    ir, jc, val = d.indices, d.indptr, d.data
    degree = np.diff(jc)
    width = (int(degree.max(initial=0)) + 1) * 2
    n_data = data.shape[0]
    out = np.zeros((n_data + width, nn))
    out[:n_data] = data
    for node in range(nn):
        k = n_data
        for i in range(jc[node], jc[node + 1]):
            out[k, node] = ir[i]
            if metric == 1:
                out[k + 1, node] = val[i]
            elif metric == 2:
                out[k + 1, node] = 1.0 / (val[i] * val[i])
            else:
                out[k + 1, node] = degree[ir[i]]
            k += 2
        out[k, node] = -1.0
    return out


def add_spatial_coupling_to_ldata(x, y, cutoff, ldata=None, min_dist=None,
                                  na_fail=True):
    """ldata with the neighbours within cutoff appended, distances as the
    values (R/distance.R:156-176)."""
    distance = distance_matrix(x, y, cutoff, min_dist=min_dist, na_fail=na_fail)
    if ldata is None:
        ldata = np.zeros((0, distance.shape[1]))
    else:
        ldata = np.asarray(ldata, dtype=np.float64)
        if ldata.ndim == 1:
            ldata = ldata.reshape(-1, 1)
        if ldata.shape[1] != distance.shape[1]:
            raise SpatialError("Number of nodes in 'ldata' and coordinates must match.")
    return ldata_sp(ldata, distance, 1)

"""Built-in model constructors: the S, G, E, N, ldata, gdata and v0
layouts of the reference's constructors, bound to the device plugins of
the same name.

  SIR    R/SIR.R:56-206      S = (S, I, R), ldata = (beta, gamma)
  SEIR   R/SEIR.R:55-158     S = (S, E, I, R), ldata = (beta, epsilon, gamma)
  SIS    R/SIS.R             S = (S, I), ldata = (beta, gamma)
  SISe   R/SISe.R:99-199     v = phi, ldata = end_t1..4, gdata (8)
  SISe3  R/SISe3.R:40-205    v = phi, ldata = end_t1..4, gdata (12)
  SISe_sp   R/SISe_sp.R      SISe with local spread: ldata = end_t1..4 and
                             the neighbour pairs (SimInf_ldata_sp, metric
                             1), gdata (8, coupling last)
  SISe3_sp  R/SISe3_sp.R     SISe3 with local spread, gdata (12)
  SIR_mparse                 mparse(c("S -> beta*S*I/N -> I",
                             "I -> gamma*I -> R", "N <- S+I+R"), gdata =
                             c(beta, gamma)) (R/mparse.R:647-724)

u0 is given as a dict of per-node arrays keyed by compartment name, or as
an (Nn x Nc) array, like the data.frame the R constructors take.
"""
import numpy as np

from .model import SimInfModel


def _u0_matrix(u0, compartments):
    if isinstance(u0, dict):
        cols = [np.atleast_1d(np.asarray(u0[c])) for c in compartments]
        n = max(len(c) for c in cols)
        cols = [np.broadcast_to(c, (n,)) for c in cols]
        return np.vstack(cols).astype(np.int64)
    u0 = np.asarray(u0)
    if u0.ndim == 1:
        u0 = u0.reshape(1, -1)
    return u0.T.astype(np.int64)


def _rep(x, n):
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    return np.resize(x, n) if len(x) != n else x


def _tile_reps(u0, v0, replicates):
    if replicates > 1:
        u0 = np.tile(u0, (1, replicates))
        if v0 is not None:
            v0 = np.tile(v0, (1, replicates))
    return u0, v0


# E matrices (column-major literals of the reference constructors).
def select_matrix_SIR():
    return np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 1], float).reshape(4, 3).T


def select_matrix_SEIR():
    return np.array([1, 0, 0, 0, 1, 1, 1, 1], float).reshape(2, 4).T


def select_matrix_SIS():
    return np.array([1, 0, 1, 1], float).reshape(2, 2).T


def select_matrix_SISe3():
    return np.array([1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0,
                     1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1],
                    float).reshape(6, 6).T


def shift_matrix_SISe3():
    return np.array([2, 2, 0, 0, 0, 0, 0, 0, 2, 2, 0, 0, 0, -1, 0, -1, 0, -1],
                    np.int32).reshape(3, 6).T


def SIR(u0, tspan, events=None, beta=None, gamma=None, replicates=1):
    u = _u0_matrix(u0, ("S", "I", "R"))
    Nn = u.shape[1]
    ldata = np.vstack([_rep(beta, Nn), _rep(gamma, Nn)])
    G = np.ones((2, 2))
    S = np.array([[-1, 0], [1, -1], [0, 1]], float)
    u, _ = _tile_reps(u, None, replicates)
    return SimInfModel("SIR", G, S, tspan, u, ldata=ldata, E=select_matrix_SIR(),
                       events=events, replicates=replicates,
                       compartments=("S", "I", "R"), ldata_names=("beta", "gamma"))


def SEIR(u0, tspan, events=None, beta=None, epsilon=None, gamma=None,
         replicates=1):
    u = _u0_matrix(u0, ("S", "E", "I", "R"))
    Nn = u.shape[1]
    ldata = np.vstack([_rep(beta, Nn), _rep(epsilon, Nn), _rep(gamma, Nn)])
    G = np.ones((3, 3))
    S = np.array([[-1, 0, 0], [1, -1, 0], [0, 1, -1], [0, 0, 1]], float)
    u, _ = _tile_reps(u, None, replicates)
    return SimInfModel("SEIR", G, S, tspan, u, ldata=ldata,
                       E=select_matrix_SEIR(), events=events,
                       replicates=replicates,
                       compartments=("S", "E", "I", "R"),
                       ldata_names=("beta", "epsilon", "gamma"))


def SIS(u0, tspan, events=None, beta=None, gamma=None, replicates=1):
    u = _u0_matrix(u0, ("S", "I"))
    Nn = u.shape[1]
    ldata = np.vstack([_rep(beta, Nn), _rep(gamma, Nn)])
    G = np.ones((2, 2))
    S = np.array([[-1, 1], [1, -1]], float)
    u, _ = _tile_reps(u, None, replicates)
    return SimInfModel("SIS", G, S, tspan, u, ldata=ldata, E=select_matrix_SIS(),
                       events=events, replicates=replicates,
                       compartments=("S", "I"), ldata_names=("beta", "gamma"))


def SISe(u0, tspan, events=None, phi=0.0, upsilon=None, gamma=None, alpha=None,
         beta_t1=None, beta_t2=None, beta_t3=None, beta_t4=None, end_t1=None,
         end_t2=None, end_t3=None, end_t4=None, epsilon=None, replicates=1):
    u = _u0_matrix(u0, ("S", "I"))
    Nn = u.shape[1]
    v0 = _rep(phi, Nn).reshape(1, -1)
    ldata = np.vstack([_rep(end_t1, Nn), _rep(end_t2, Nn), _rep(end_t3, Nn),
                       _rep(end_t4, Nn)])
    gdata = np.array([upsilon, gamma, alpha, beta_t1, beta_t2, beta_t3,
                      beta_t4, epsilon], float)
    G = np.ones((2, 2))
    S = np.array([[-1, 1], [1, -1]], float)
    u, v0 = _tile_reps(u, v0, replicates)
    return SimInfModel("SISe", G, S, tspan, u, ldata=ldata, gdata=gdata, v0=v0,
                       E=select_matrix_SIS(), events=events,
                       replicates=replicates,
                       compartments=("S", "I"), v0_names=("phi",),
                       ldata_names=("end_t1", "end_t2", "end_t3", "end_t4"),
                       gdata_names=("upsilon", "gamma", "alpha", "beta_t1", "beta_t2",
                                    "beta_t3", "beta_t4", "epsilon"))


def _sise3_G_S():
    """G and S of SISe3 and SISe3_sp (R/SISe3.R:162-178)."""
    G = np.array([1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0,
                  0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1],
                 float).reshape(6, 6).T
    S = np.array([-1, 1, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 0, 0, -1, 1, 0, 0,
                  0, 0, 1, -1, 0, 0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, 1, -1],
                 float).reshape(6, 6).T
    return G, S


def SISe3(u0, tspan, events=None, phi=0.0, upsilon_1=None, upsilon_2=None,
          upsilon_3=None, gamma_1=None, gamma_2=None, gamma_3=None, alpha=None,
          beta_t1=None, beta_t2=None, beta_t3=None, beta_t4=None, end_t1=None,
          end_t2=None, end_t3=None, end_t4=None, epsilon=None, replicates=1):
    u = _u0_matrix(u0, ("S_1", "I_1", "S_2", "I_2", "S_3", "I_3"))
    Nn = u.shape[1]
    v0 = _rep(phi, Nn).reshape(1, -1)
    ldata = np.vstack([_rep(end_t1, Nn), _rep(end_t2, Nn), _rep(end_t3, Nn),
                       _rep(end_t4, Nn)])
    gdata = np.array([upsilon_1, upsilon_2, upsilon_3, gamma_1, gamma_2,
                      gamma_3, alpha, beta_t1, beta_t2, beta_t3, beta_t4,
                      epsilon], float)
    G, S = _sise3_G_S()
    u, v0 = _tile_reps(u, v0, replicates)
    return SimInfModel("SISe3", G, S, tspan, u, ldata=ldata, gdata=gdata, v0=v0,
                       E=select_matrix_SISe3(), N=shift_matrix_SISe3(),
                       events=events, replicates=replicates,
                       compartments=("S_1", "I_1", "S_2", "I_2", "S_3", "I_3"),
                       v0_names=("phi",),
                       ldata_names=("end_t1", "end_t2", "end_t3", "end_t4"),
                       gdata_names=("upsilon_1", "upsilon_2", "upsilon_3", "gamma_1",
                                    "gamma_2", "gamma_3", "alpha", "beta_t1", "beta_t2",
                                    "beta_t3", "beta_t4", "epsilon"))


def _check_distance(distance, Nn):
    """check_distance_matrix (R/check_arguments.R:285-296)."""
    import scipy.sparse as sp
    if distance is None:
        raise ValueError("'distance' is missing.")
    if not sp.issparse(distance):
        raise ValueError("The 'distance' argument must be of type 'dgCMatrix'.")
    if (distance.data < 0).any():
        raise ValueError("All values in the 'distance' matrix must be >= 0.")
    if distance.shape != (Nn, Nn):
        raise ValueError("The number of nodes in 'data' and 'distance' are not equal.")


def _sp_ldata(Nn, end_t, distance):
    from .spatial import ldata_sp
    _check_distance(distance, Nn)
    ends = np.vstack([_rep(e, Nn) for e in end_t])
    ldata = ldata_sp(ends, distance, 1)
    names = ["end_t1", "end_t2", "end_t3", "end_t4"] + \
        ["neighbor_%d" % k for k in range(ldata.shape[0] - 4)]
    return ldata, names


def SISe_sp(u0, tspan, events=None, phi=0.0, upsilon=None, gamma=None,
            alpha=None, beta_t1=None, beta_t2=None, beta_t3=None, beta_t4=None,
            end_t1=None, end_t2=None, end_t3=None, end_t4=None, coupling=None,
            distance=None, replicates=1):
    """SISe with local spread of phi among the nodes within the distance
    matrix (R/SISe_sp.R:99-169); distance as from spatial.distance_matrix."""
    u = _u0_matrix(u0, ("S", "I"))
    Nn = u.shape[1]
    v0 = _rep(phi, Nn).reshape(1, -1)
    ldata, ld_names = _sp_ldata(Nn, (end_t1, end_t2, end_t3, end_t4), distance)
    gdata = np.array([upsilon, gamma, alpha, beta_t1, beta_t2, beta_t3,
                      beta_t4, coupling], float)
    G = np.ones((2, 2))
    S = np.array([[-1, 1], [1, -1]], float)
    u, v0 = _tile_reps(u, v0, replicates)
    return SimInfModel("SISe_sp", G, S, tspan, u, ldata=ldata, gdata=gdata, v0=v0,
                       E=select_matrix_SIS(), events=events,
                       replicates=replicates,
                       compartments=("S", "I"), v0_names=("phi",),
                       ldata_names=ld_names,
                       gdata_names=("upsilon", "gamma", "alpha", "beta_t1", "beta_t2",
                                    "beta_t3", "beta_t4", "coupling"))


def SISe3_sp(u0, tspan, events=None, phi=0.0, upsilon_1=None, upsilon_2=None,
             upsilon_3=None, gamma_1=None, gamma_2=None, gamma_3=None,
             alpha=None, beta_t1=None, beta_t2=None, beta_t3=None, beta_t4=None,
             end_t1=None, end_t2=None, end_t3=None, end_t4=None, distance=None,
             coupling=None, replicates=1):
    """SISe3 with local spread of phi (R/SISe3_sp.R:112-205)."""
    u = _u0_matrix(u0, ("S_1", "I_1", "S_2", "I_2", "S_3", "I_3"))
    Nn = u.shape[1]
    v0 = _rep(phi, Nn).reshape(1, -1)
    ldata, ld_names = _sp_ldata(Nn, (end_t1, end_t2, end_t3, end_t4), distance)
    gdata = np.array([upsilon_1, upsilon_2, upsilon_3, gamma_1, gamma_2,
                      gamma_3, alpha, beta_t1, beta_t2, beta_t3, beta_t4,
                      coupling], float)
    G, S = _sise3_G_S()
    u, v0 = _tile_reps(u, v0, replicates)
    return SimInfModel("SISe3_sp", G, S, tspan, u, ldata=ldata,
                       gdata=gdata, v0=v0, E=select_matrix_SISe3(),
                       N=shift_matrix_SISe3(), events=events,
                       replicates=replicates,
                       compartments=("S_1", "I_1", "S_2", "I_2", "S_3", "I_3"),
                       v0_names=("phi",), ldata_names=ld_names,
                       gdata_names=("upsilon_1", "upsilon_2", "upsilon_3", "gamma_1",
                                    "gamma_2", "gamma_3", "alpha", "beta_t1", "beta_t2",
                                    "beta_t3", "beta_t4", "coupling"))


def SIR_mparse(u0, tspan, beta, gamma, events=None, E=None, N=None,
               replicates=1):
    """The model mparse builds for c("S -> beta*S*I/N -> I",
    "I -> gamma*I -> R", "N <- S+I+R") with gdata c(beta, gamma):
    S = dest - from per transition (R/mparse.R:234-236) and
    G = (depends %*% abs(S)) > 0 (R/mparse.R:485-491), which is full here
    because the variable N makes trFun1 depend on S, I and R."""
    u = _u0_matrix(u0, ("S", "I", "R"))
    S = np.array([[-1, 0], [1, -1], [0, 1]], float)
    depends = np.array([[1, 1, 1], [0, 1, 0]], float)
    G = ((depends @ np.abs(S)) > 0).astype(float)
    u, _ = _tile_reps(u, None, replicates)
    return SimInfModel("SIR_mparse", G, S, tspan, u, gdata=[beta, gamma],
                       E=E if E is not None else np.zeros((3, 0)), N=N,
                       events=events, replicates=replicates,
                       compartments=("S", "I", "R"), gdata_names=("beta", "gamma"))

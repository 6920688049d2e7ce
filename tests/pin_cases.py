"""Models whose laws are fixed by the mathematics, and the reference's
event and post-time-step error cases, shared by the checker's CPU tests
(tests/test_model_pins.py) and the GPU tests (tests/test_gpu_pins.py).

The expected laws are written here from the model definitions in the
reference (src/models/SEIR.c:43-112, src/models/SISe3.c:51-242, the
interval table documented at src/misc/SimInf_forward_euler_linear_decay.c:
26-55), not from the plugin code under test: a transcription slip in a
plugin body (a wrong compartment, parameter index or sign) changes one of
these laws.
"""
import numpy as np
import scipy.linalg
import scipy.sparse as sp
from scipy import stats

import siminf_amd as si


def chisq_p(obs, pmf, n, min_exp=10):
    """Chi-square p-value with the bins of expected count < min_exp
    pooled into one."""
    obs = np.asarray(obs, float)
    e = np.asarray(pmf, float) * n
    keep = e >= min_exp
    o, ex = list(obs[keep]), list(e[keep])
    if e[~keep].sum() > 0 or obs[~keep].sum() > 0:
        o.append(obs[~keep].sum())
        ex.append(e[~keep].sum())
    o, ex = np.array(o), np.array(ex)
    if ex[-1] < 5 and len(ex) > 1:
        o[-2] += o[-1]
        ex[-2] += ex[-1]
        o, ex = o[:-1], ex[:-1]
    ex *= o.sum() / ex.sum()
    return stats.chisquare(o, ex).pvalue


# ------------------------------------------------------------- SEIR ---

SEIR_PARS = dict(beta=0.9, epsilon=0.5, gamma=0.4)
SEIR_START = (3, 1, 2, 0)


def seir_one_day_law(start=SEIR_START, beta=0.9, epsilon=0.5, gamma=0.4, t=1.0):
    """Law of (S, E, I, R) after time t for one SEIR node: the matrix
    exponential of the generator of S -> E at beta*S*I/n (n = S+E+I+R),
    E -> I at epsilon*E, I -> R at gamma*I (SEIR.c:43-112)."""
    N = sum(start)
    states = [(s, e, i, N - s - e - i) for s in range(N + 1) for e in range(N + 1 - s)
              for i in range(N + 1 - s - e)]
    index = {x: k for k, x in enumerate(states)}
    Q = np.zeros((len(states), len(states)))
    for (s, e, i, r), k in index.items():
        moves = [((s - 1, e + 1, i, r), beta * s * i / N),
                 ((s, e - 1, i + 1, r), epsilon * e),
                 ((s, e, i - 1, r + 1), gamma * i)]
        for dst, rate in moves:
            if rate > 0:
                Q[k, index[dst]] += rate
                Q[k, k] -= rate
    return states, index, scipy.linalg.expm(Q * t)[index[tuple(start)]]


def seir_brute_force_model(n_nodes):
    """n_nodes independent copies of the one-node chain; column 0 of U
    holds the state after the first day (tspan = 0)."""
    s, e, i, r = SEIR_START
    return si.SEIR(u0={"S": np.full(n_nodes, s), "E": e, "I": i, "R": r},
                   tspan=np.array([0.0]), **SEIR_PARS)


def seir_brute_force_p(U, n_nodes):
    """Chi-square p-value of the nodes' states in column 0 of U."""
    states, index, P = seir_one_day_law()
    X = U[:, 0].reshape(n_nodes, 4)
    obs = np.zeros(len(states))
    for row in X:
        obs[index[tuple(int(x) for x in row)]] += 1
    return chisq_p(obs, P, n_nodes)


def seir_pure_death_model(n_nodes):
    """No infection (beta = 0): E0 = 8 exposed decay to I at rate epsilon
    and I0 = 6 infected to R at rate gamma, so that after t days the E
    count is Binomial(8, exp(-epsilon t)) and the R count of a node that
    starts with I only is Binomial(6, 1 - exp(-gamma t))."""
    half = n_nodes // 2
    E = np.r_[np.full(half, 8), np.zeros(n_nodes - half, int)]
    I = np.r_[np.zeros(half, int), np.full(n_nodes - half, 6)]
    return si.SEIR(u0={"S": 5, "E": E, "I": I, "R": 0}, tspan=np.array([0.0, 2.0]),
                   beta=0.0, epsilon=0.35, gamma=0.2)


def seir_pure_death_ps(U, n_nodes):
    half = n_nodes // 2
    X = U.reshape(n_nodes, 4, -1)
    ps = []
    for col, t in ((0, 1.0), (1, 3.0)):
        # S never moves without infection.
        assert (X[:, 0, col] == 5).all()
        E = X[:half, 1, col]
        ps.append(chisq_p(np.bincount(E, minlength=9),
                          stats.binom.pmf(np.arange(9), 8, np.exp(-0.35 * t)), half))
        R = X[half:, 3, col]
        ps.append(chisq_p(np.bincount(R, minlength=7),
                          stats.binom.pmf(np.arange(7), 6, 1 - np.exp(-0.2 * t)),
                          n_nodes - half))
        # The first half's E + I + R and the second half's I + R keep their
        # totals.
        assert (X[:half, 1:, col].sum(1) == 8).all()
        assert (X[half:, 2:, col].sum(1) == 6).all()
    return ps


# ------------------------------------------------------------ SISe3 ---

SISE3_GDATA_NAMES = ("upsilon_1", "upsilon_2", "upsilon_3", "gamma_1", "gamma_2",
                     "gamma_3", "alpha", "beta_t1", "beta_t2", "beta_t3", "beta_t4",
                     "epsilon")
SISE3_START = (5, 4, 6, 3, 7, 2)   # S_1, I_1, S_2, I_2, S_3, I_3


def sise3_model(n_nodes, tspan, u=SISE3_START, phi=1.0, ends=(91, 182, 273, 365),
                **pars):
    """SISe3 with every rate parameter 0 unless given."""
    kw = {k: 0.0 for k in SISE3_GDATA_NAMES}
    kw.update(pars)
    names = ("S_1", "I_1", "S_2", "I_2", "S_3", "I_3")
    u0 = {nm: np.full(n_nodes, x) if np.ndim(x) == 0 else np.asarray(x)
          for nm, x in zip(names, u)}
    return si.SISe3(u0=u0, tspan=np.asarray(tspan, float), phi=phi,
                    end_t1=ends[0], end_t2=ends[1], end_t3=ends[2], end_t4=ends[3],
                    **kw)


# One transition at a time: (parameter, moving compartment, receiving
# compartment, rate per individual with phi = 0.7).  With alpha = epsilon
# = beta_t = 0 the pressure phi stays at its initial value, so the moving
# count after t days is Binomial(n0, 1 - exp(-rate t)) and every other
# compartment keeps its value (SISe3.c:51-169: S_k -> I_k at
# upsilon_k*phi*S_k, I_k -> S_k at gamma_k*I_k).
SISE3_SINGLE = [
    ("upsilon_1", 0, 1, 0.3 * 0.7),
    ("gamma_1", 1, 0, 0.25),
    ("upsilon_2", 2, 3, 0.2 * 0.7),
    ("gamma_2", 3, 2, 0.35),
    ("upsilon_3", 4, 5, 0.15 * 0.7),
    ("gamma_3", 5, 4, 0.45),
]
SISE3_SINGLE_VALUE = {"upsilon_1": 0.3, "gamma_1": 0.25, "upsilon_2": 0.2,
                      "gamma_2": 0.35, "upsilon_3": 0.15, "gamma_3": 0.45}


def sise3_single_model(param, n_nodes):
    return sise3_model(n_nodes, [0.0, 1.0, 2.0], phi=0.7,
                       **{param: SISE3_SINGLE_VALUE[param]})


def sise3_single_ps(U, V, n_nodes, param):
    """p-values of the moving count per column; asserts that the other
    compartments and phi never change."""
    _, src, dst, rate = next(c for c in SISE3_SINGLE if c[0] == param)
    X = U.reshape(n_nodes, 6, -1)
    n0 = SISE3_START[src]
    ps = []
    for col in range(X.shape[2]):
        t = col + 1.0
        for c in range(6):
            if c not in (src, dst):
                assert (X[:, c, col] == SISE3_START[c]).all(), (param, c)
        assert (X[:, src, col] + X[:, dst, col] == n0 + SISE3_START[dst]).all()
        moved = n0 - X[:, src, col]
        ps.append(chisq_p(np.bincount(moved, minlength=n0 + 1),
                          stats.binom.pmf(np.arange(n0 + 1), n0, 1 - np.exp(-rate * t)),
                          n_nodes))
    assert (V == 0.7).all()
    return ps


SISE3_CHAIN_PARS = dict(upsilon_1=0.3, upsilon_2=0.2, upsilon_3=0.15,
                        gamma_1=0.25, gamma_2=0.35, gamma_3=0.45)


def sis_chain_law(S0, I0, up, gam, t=1.0):
    """Law of I after t for an SIS chain S -> I at up*S, I -> S at gam*I
    (phi folded into up)."""
    N = S0 + I0
    Q = np.zeros((N + 1, N + 1))
    for i in range(N + 1):
        s = N - i
        if s > 0:
            Q[i, i + 1] += up * s
            Q[i, i] -= up * s
        if i > 0:
            Q[i, i - 1] += gam * i
            Q[i, i] -= gam * i
    return scipy.linalg.expm(Q * t)[I0]


def sise3_chain_model(n_nodes):
    """All six transitions at once with phi held at 0.7: the three age
    classes are independent SIS chains."""
    return sise3_model(n_nodes, [0.0], phi=0.7, **SISE3_CHAIN_PARS)


def sise3_chain_ps(U, n_nodes):
    X = U[:, 0].reshape(n_nodes, 6)
    ps = []
    for k in range(3):
        S0, I0 = SISE3_START[2 * k], SISE3_START[2 * k + 1]
        up = SISE3_CHAIN_PARS["upsilon_%d" % (k + 1)] * 0.7
        gam = SISE3_CHAIN_PARS["gamma_%d" % (k + 1)]
        assert (X[:, 2 * k] + X[:, 2 * k + 1] == S0 + I0).all()
        ps.append(chisq_p(np.bincount(X[:, 2 * k + 1], minlength=S0 + I0 + 1),
                          sis_chain_law(S0, I0, up, gam), n_nodes))
    return ps


def decay_interval(day, ends):
    """The interval (1..4) of day 0 <= day < 365 in the three cases of the
    table at SimInf_forward_euler_linear_decay.c:30-42."""
    e1, e2, e3, e4 = ends
    if e1 < e2 < e3 < e4:
        table = [(0, e1, 1), (e1, e2, 2), (e2, e3, 3), (e3, e4, 4), (e4, 365, 1)]
    elif e3 < e4 < e1 < e2:
        table = [(0, e3, 3), (e3, e4, 4), (e4, e1, 1), (e1, e2, 2), (e2, 365, 3)]
    elif e4 < e1 < e2 < e3:
        table = [(0, e4, 4), (e4, e1, 1), (e1, e2, 2), (e2, e3, 3), (e3, 365, 4)]
    else:
        raise ValueError("not one of the documented cases")
    for lo, hi, k in table:
        if lo <= day < hi:
            return k
    raise AssertionError(day)


PHI_CASES = [(91, 182, 273, 365), (120, 200, 30, 80), (100, 180, 300, 40)]
PHI_U = [(10, 1, 20, 2, 30, 3), (0, 0, 0, 0, 0, 0), (7, 0, 0, 5, 1, 0)]
PHI_PARS = dict(alpha=0.7, epsilon=0.01, beta_t1=0.1, beta_t2=0.2, beta_t3=0.3,
                beta_t4=0.4)


def phi_model(ends):
    """No transitions (upsilon = gamma = 0): u is fixed and phi follows
    phi <- phi (1 - beta(day)) + alpha I/n + epsilon after every day, with
    I = I_1 + I_2 + I_3 and n the node's total, epsilon alone for an empty
    node (SISe3.c:203-242), beta(day) the interval's value, day = t mod
    365 for the day's start time t.  400 days cross a year's end."""
    u = [np.array([PHI_U[k][c] for k in range(len(PHI_U))]) for c in range(6)]
    return sise3_model(len(PHI_U), np.arange(400.0), u=u, phi=0.5, ends=ends,
                       **PHI_PARS)


def phi_expected(ends):
    """V (nodes x 400 columns): column j is phi after the post-time-step
    call of day j (time j)."""
    beta = [None, PHI_PARS["beta_t1"], PHI_PARS["beta_t2"], PHI_PARS["beta_t3"],
            PHI_PARS["beta_t4"]]
    out = np.zeros((len(PHI_U), 400))
    for k, u in enumerate(PHI_U):
        I = u[1] + u[3] + u[5]
        n = I + u[0] + u[2] + u[4]
        phi = 0.5
        for j in range(400):
            phi = phi * (1.0 - beta[decay_interval(j % 365, ends)])
            phi = phi + (PHI_PARS["alpha"] * I / n + PHI_PARS["epsilon"]
                         if n > 0 else PHI_PARS["epsilon"])
            out[k, j] = phi
    return out


# --------------------------------------------------- reference errors ---

def _ev(event, time, node, dest, n, proportion, select, shift):
    return {"event": [event], "time": [time], "node": [node], "dest": [dest], "n": [n],
            "proportion": [proportion], "select": [select], "shift": [shift]}


def _sir_two_nodes(events):
    return si.SIR(u0={"S": [10, 10], "I": [0, 0], "R": [0, 0]}, tspan=[1.0, 2.0, 3.0],
                  events=events, beta=0.16, gamma=0.077)


def _sir_select_four(event, n=1, proportion=0.0, shift=0, col4=(0, 0, 0), N=None):
    """tests/sample_select.R:857-941: one node (1, 1, 1), one event at
    time 1 with select 4, E's fourth column replaced."""
    m = si.SIR(u0={"S": [1], "I": [1], "R": [1]}, tspan=[1.0],
               events=_ev(event, 1, 1, 0, n, proportion, 4, shift), beta=0.0, gamma=0.0)
    E = np.array([[1, 0, 0, col4[0]], [0, 1, 0, col4[1]], [0, 0, 1, col4[2]]], float)
    m.E = m.events["E"] = sp.csc_matrix(E)
    m.E.sort_indices()
    if N is not None:
        m.N = m.events["N"] = np.asarray(N, dtype=np.int32).reshape(3, -1, order="F")
    return m


def _set(m, key, value):
    m.events[key][:] = value
    return m


def _sise3_ref(**override):
    """tests/SISe3.R:1395-1445: phi = 1, the reference's parameters,
    tspan 0..9."""
    pars = dict(upsilon_1=0.0357, upsilon_2=0.0357, upsilon_3=0.00935, gamma_1=0.1,
                gamma_2=0.1, gamma_3=0.1, alpha=1.0, beta_t1=0.19, beta_t2=0.085,
                beta_t3=0.075, beta_t4=0.185, epsilon=0.000011)
    pars.update(override)
    u = [np.array([0, 5, 10, 20, 40, 1]), np.array([0, 0, 1, 2, 0, 0]),
         np.array([3, 5, 0, 7, 8, 2]), np.array([0, 1, 0, 0, 0, 0]),
         np.array([9, 0, 4, 11, 6, 3]), np.array([0, 0, 0, 1, 2, 0])]
    return sise3_model(6, np.arange(10.0), u=u, phi=1.0, ends=(91, 182, 273, 365),
                       **pars)


def error_cases():
    """(name, model builder, expected code): the reference's event and
    post-time-step error tests restated, and one known answer."""
    return [
        # tests/SimInf_events.R:537-545
        ("dest_out_of_bounds", lambda: _sir_two_nodes(_ev(3, 2, 1, 3, 1, 0.0, 2, 0)), -13),
        # tests/SimInf_events.R:547-555 (node set to -1 after validation)
        ("node_out_of_bounds",
         lambda: _set(_sir_two_nodes(_ev(3, 2, 1, 2, 1, 0.0, 2, 0)), "node", -1), -14),
        # tests/SimInf_events.R:557-565 (event type set to 4)
        ("undefined_event",
         lambda: _set(_sir_two_nodes(_ev(0, 2, 1, 0, 1, 0.0, 2, 0)), "event", 4), -3),
        # tests/sample_select.R:859-885: an empty E column
        ("exit_empty_select", lambda: _sir_select_four(0), -8),
        ("enter_empty_select", lambda: _sir_select_four(1), -8),
        ("enter_negative_n", lambda: _sir_select_four(1, n=-1), -8),
        # tests/sample_select.R:887-904
        ("enter_proportion_below_0",
         lambda: _sir_select_four(1, n=0, proportion=-1.0, col4=(1, 1, 1)), -18),
        ("enter_proportion_above_1",
         lambda: _sir_select_four(1, n=0, proportion=1.1, col4=(1, 1, 1)), -18),
        # tests/sample_select.R:913-919: a shift without N
        ("enter_shift_without_N",
         lambda: _set(_sir_select_four(1, shift=1, col4=(1, 1, 1)), "select", 1), -15),
        # tests/sample_select.R:921-932
        ("enter_shift_out_of_bounds_high",
         lambda: _set(_sir_select_four(1, shift=1, col4=(1, 1, 1), N=[3, 0, 0]),
                      "select", 1), -17),
        ("enter_shift_out_of_bounds_low",
         lambda: _set(_sir_select_four(1, shift=1, col4=(1, 1, 1), N=[-1, 0, 0]),
                      "select", 1), -17),
        # SimInf_solver.c:764-782: an internal transfer needs N and a shift
        ("internal_without_N", lambda: _sir_select_four(2, shift=1, col4=(1, 0, 0)), -15),
        ("internal_shift_invalid",
         lambda: _sir_select_four(2, shift=0, col4=(1, 0, 0), N=[1, 0, 0]), -16),
        ("internal_shift_out_of_bounds",
         lambda: _sir_select_four(2, shift=1, col4=(1, 0, 0), N=[5, 0, 0]), -17),
        # SimInf_solver.c:858-868: an external transfer shifting without N
        ("external_shift_without_N",
         lambda: _sir_two_nodes(_ev(3, 2, 1, 2, 1, 0.0, 1, 1)), -15),
        # tests/SISe3.R:1395-1445
        ("sise3_v_not_finite",
         lambda: _sise3_ref(beta_t1=np.inf, beta_t2=np.inf, beta_t3=np.inf,
                            beta_t4=np.inf), -7),
        ("sise3_v_negative", lambda: _sise3_ref(epsilon=-10.000011), -10),
    ]


def enter_shift_known_answer():
    """tests/sample_select.R:934-940: an enter event of one individual
    into S shifted by N = (1, 0, 0) lands in I: U = (1, 2, 1)."""
    m = _set(_sir_select_four(1, shift=1, col4=(1, 1, 1), N=[1, 0, 0]), "select", 1)
    return m, np.array([[1], [2], [1]])

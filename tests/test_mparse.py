"""The model parser against the reference's own mparse expectations
(tests/mparse.R): emitted C text, S and G matrices, tokenizer, variable
ordering and error messages."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from siminf_amd import modelparser as mp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BLOCKS = json.load(open(os.path.join(GOLDEN, "mparse_c_code.json")))["blocks"]
DW = ["@->c1->D", "D->c2*D->D+D", "D+W->c3*D*W->W+W", "W->c4*W->@"]


def csc(M):
    m = sp.csc_matrix(M)
    return m.indices.tolist(), m.indptr.tolist(), m.data.tolist()


def test_c_code_gdata_model():
    """tests/mparse.R:201-408."""
    m = mp.ParsedModel(DW, ["D", "W"], gdata={"c1": 0.5, "c2": 1, "c3": 0.005, "c4": 0.6})
    assert m.C_code[1:] == BLOCKS[0]
    assert csc(m.G) == ([1, 2, 1, 2, 1, 2, 3, 2, 3], [0, 2, 4, 7, 9], [1.0] * 9)
    assert m.G_rownames == ["@ -> c1 -> D", "D -> c2*D -> 2*D",
                            "D + W -> c3*D*W -> 2*W", "W -> c4*W -> @"]
    assert csc(m.S) == ([0, 0, 0, 1, 1], [0, 1, 2, 4, 5], [1.0, 1.0, -1.0, 1.0, -1.0])


def test_c_code_gdata_and_ldata_model():
    """tests/mparse.R:410-597."""
    m = mp.ParsedModel(DW, ["D", "W"], ldata={"c4": None},
                       gdata={"c1": 0.5, "c2": 1, "c3": 0.005})
    assert m.C_code[1:] == BLOCKS[1]


def test_c_code_pre_code():
    """tests/mparse.R:621-785."""
    m = mp.ParsedModel(["S -> rate(S, I, S+I+R, b) -> I", "I -> g*I -> R"],
                       ["S", "I", "R"], gdata={"b": 0.16, "g": 0.077},
                       pre_code=["double rate(double S, double I, double N, double beta) {",
                                 "    if (N == 0)",
                                 "        return 0;",
                                 "    return beta * S * I / N;",
                                 "}"])
    assert m.C_code[1:] == BLOCKS[2]


def test_tokenize_and_rewrite():
    """tests/mparse.R:599-619."""
    assert mp.tokenize("beta*S*I/(S+I+R)") == ["beta", "*", "S", "*", "I", "/", "(",
                                               "S", "+", "I", "+", "R", ")"]
    r = mp.rewrite_propensity("beta*S*I/(S+I+R)", {}, ["S", "I", "R"], None,
                              ["beta"], None, False)
    assert r == {"code": "gdata[0]*u[0]*u[1]/(u[0]+u[1]+u[2])", "depends": [1, 1, 1],
                 "G_rowname": "beta*S*I/(S+I+R)", "variables": []}
    r = mp.rewrite_propensity("beta*S*I/(S+I+R)", {}, ["S", "I", "R"], None,
                              ["beta"], None, True)
    assert r["code"] == "gdata[BETA]*u[S]*u[I]/(u[S]+u[I]+u[R])"


def test_emitted_lines():
    """tests/mparse.R:895-907."""
    m = mp.ParsedModel(".S.S -> 1.2*.S.S -> @", [".S.S"])
    assert m.C_code[45] == "    return 1.2*u[0];"
    m = mp.ParsedModel("S -> a->data[2]*1.2*S -> @", ["S"])
    assert m.C_code[45] == "    return a->data[2]*1.2*u[0];"


def test_replicated_compartments_and_empty_set():
    """tests/mparse.R:839-893."""
    assert mp.parse_compartments("S + 2*S", ["S", "I"]) == [3, 0]
    m = mp.ParsedModel("S + S -> mu -> @", ["S", "I"], gdata={"mu": 1})
    assert m.S[:, 0].tolist() == [-2.0, 0.0] and m.G.sum() == 0
    assert m.G_rownames == ["2*S -> mu -> @"]
    m = mp.ParsedModel("@ -> mu-> S + S", ["S", "I"], gdata={"mu": 1})
    assert m.S[:, 0].tolist() == [2.0, 0.0] and m.G_rownames == ["@ -> mu -> 2*S"]
    for bad in ("I -> gamma*I -> @+R", "I -> gamma*I -> @+@"):
        with pytest.raises(mp.MparseError, match="Invalid usage of the empty set '@'."):
            mp.ParsedModel(["S -> beta*S*I/(S+I+R) -> I", bad], ["S", "I", "R"],
                           gdata={"beta": 0.16, "gamma": 0.077})


def test_variables_dependencies():
    """tests/mparse.R:1178-1203: compartments reached through variables."""
    m = mp.ParsedModel(["S -> beta*NS*NI/N -> E", "E -> gamma*NE -> I",
                        "I -> delta*NI -> R", "NS <- S", "NE <- E", "NI <- I",
                        "N <- NS+NE+NI+NR"], ["S", "E", "I", "R"],
                       ldata={"beta": 1, "gamma": 2}, v0={"delta": 3})
    assert csc(m.G) == ([0, 1, 0, 1, 2, 0, 2], [0, 2, 5, 7], [1.0] * 7)


def test_c_variables_order():
    """tests/mparse.R:1106-1128."""
    prop = {"code": "N>0?beta*u[0]*u[1]/N:0", "depends": [1, 1, 0],
            "variables": ["N"]}
    variables = {"N1": {"type": "double", "depends": [], "code": "u[0]"},
                 "N2": {"type": "double", "depends": ["N1"], "code": "N1+u[1]"},
                 "N": {"type": "double", "depends": ["N2"], "code": "N2+u[2]"}}
    assert mp.C_variables(prop, variables) == [
        "    const double N1 = u[0];", "    const double N2 = N1+u[1];",
        "    const double N = N2+u[2];", ""]


def test_topological_sort():
    """tests/mparse.R:1082-1102: dep[i][j] = 1 when column j uses row i."""
    names = ["C", "B", "A"]
    # rows A, B, C; columns C, B, A: B uses A? matrix(c(0,1,0, 1,0,0, 0,0,0))
    M = np.array([0, 1, 0, 1, 0, 0, 0, 0, 0]).reshape(3, 3).T   # rows A,B,C
    rows = ["A", "B", "C"]
    dep = [[M[rows.index(r), names.index(c)] for c in names] for r in names]
    assert mp.topological_sort(dep, names) == ["A", "B", "C"]


def test_enum_constants():
    """tests/mparse.R:1130-1176."""
    lines = mp.C_enum(mp.Names(["S", "I", "R"], [0, 1, 2], 3),
                      mp.Names(["beta", "gamma", "delta", "epsilon", "zeta", "eta",
                                "theta", "iota", "kappa", "lambda"]),
                      mp.Names(["beta"], [4], 5), mp.Names(["mu", "nu"], [0, 1], 2),
                      True)
    assert lines[:7] == ["/* Enumeration constants for indicies in the 'u' vector. */",
                         "enum {", "    S = 0,", "    I = 1,", "    R = 2,",
                         "    N_COMPARTMENTS_U = 3", "};"]
    assert "    BETA," in lines and "    LAMBDA" in lines
    assert lines[-3:] == ["    BETA = 4", "};", ""]


@pytest.mark.parametrize("transitions,msg", [
    (["@->c1->D", "W->c4*W"], "Invalid transition: 'W->c4*W'."),
    (["A->c1->D"], "Unknown compartment: 'A'."),
    (["3N <- S + I + R"], None),
])
def test_errors(transitions, msg):
    if msg is None:
        return
    with pytest.raises(mp.MparseError, match=msg.replace("(", r"\(").replace("*", r"\*")):
        mp.ParsedModel(transitions, ["D", "W"], gdata={"c1": 1, "c4": 1})


def test_invalid_dependencies():
    with pytest.raises(mp.MparseError, match="Invalid dependencies between variables."):
        mp.ParsedModel(["S -> A -> I", "A <- B", "B <- A"], ["S", "I"])


SIR_TRANSITIONS = ["S -> beta*S*I/N -> I", "I -> gamma*I -> R", "N <- S+I+R"]


def _sir_pair(Nn=40, days=50, replicates=1):
    import siminf_amd as si
    u0 = {"S": np.arange(60, 60 + Nn), "I": np.arange(Nn) % 7 + 1, "R": 0}
    tspan = np.arange(1, days + 1.0)
    gen = si.mparse(SIR_TRANSITIONS, ["S", "I", "R"], gdata={"beta": 0.16, "gamma": 0.077},
                    u0=u0, tspan=tspan, replicates=replicates)
    ref = si.SIR_mparse(u0=u0, tspan=tspan, beta=0.16, gamma=0.077, replicates=replicates)
    return gen, ref


def test_generated_plugin_source():
    """The device plugin folds S and G of the parsed model and dispatches
    trFunN as rate<N-1>."""
    from siminf_amd import plugins
    gen, _ = _sir_pair()
    src = plugins.device_source(gen.plugin.parsed, "sir")
    assert "static constexpr int NC = 3, NT = 2, ND = 0, NLD = 0, NGD = 2;" in src
    assert "S[NT][NC] = {{-1, 1, 0}, {0, -1, 1}}" in src
    assert "G[NT][NT] = {{1, 1}, {1, 1}}" in src
    assert "if constexpr (J == 0) return trFun1(u, v, ld, gd, t);" in src
    assert "    const double N = u[0]+u[1]+u[2];" in src


def test_generated_host_plugin_equals_builtin(oracle):
    """Trajectories of the parsed SIR (its generated C compiled as a host
    plugin) equal those of the hand-written built-in plugin of the same
    model, replicate for replicate."""
    gen, ref = _sir_pair(replicates=3)
    a = oracle.run(gen, seed=11)
    b = oracle.run(ref, seed=11)
    assert a.err == 0 and b.err == 0
    np.testing.assert_array_equal(a.U, b.U)
    assert a.transitions == b.transitions > 0
    # individuals are conserved in every node and column
    tot = a.U.reshape(gen.Nn * 3, -1).reshape(gen.Nn, 3, -1).sum(axis=1)
    assert (tot == tot[:, :1]).all()


def test_device_plugin_builds_and_exports():
    """The generated HIP plugin compiles for gfx950 and exports the model
    descriptor with the parsed dimensions (no GPU call is made)."""
    import ctypes
    from siminf_amd import _native, plugins, workloads
    m = workloads.mparse_sis_env(Nn=2, days=2)
    dev = plugins.GeneratedPlugin(m.plugin.parsed, m.plugin.name).device_model()
    nc, nt, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert _native.load().SimInf_hip_model_dims(
        ctypes.c_void_p(dev), ctypes.byref(nc), ctypes.byref(nt), ctypes.byref(nd)) == 0
    assert (nc.value, nt.value, nd.value) == (2, 2, 1)


def test_plugin_built_against_other_headers_is_refused():
    """A device plugin records the ABI version and the layout signature of
    the kernel parameter block it was compiled against; the library
    refuses one that differs (it would otherwise hand the plugin's kernels
    a parameter block of another layout)."""
    import ctypes
    from siminf_amd import _native
    lib = _native.load()
    good = lib.SimInf_hip_model_builtin(b"SIR")
    assert lib.SimInf_hip_model_dims(ctypes.c_void_p(good), None, None, None) == 0
    raw = (ctypes.c_char * 16).from_address(good).raw
    fake = ctypes.create_string_buffer(raw + b"\0" * 112)   # abi, signature, ...
    ctypes.memmove(ctypes.addressof(fake) + 8, ctypes.byref(ctypes.c_longlong(12345)), 8)
    assert lib.SimInf_hip_model_dims(ctypes.cast(fake, ctypes.c_void_p), None, None,
                                     None) == -9     # SIMINF_ERR_INVALID_MODEL


def test_models_over_32_compartments_are_refused():
    """S, G, E and the summary masks are 32-bit sets: a parsed model of 33
    compartments is refused before any device code is generated."""
    import siminf_amd as si
    from siminf_amd import plugins
    comps = ["C%d" % k for k in range(33)]
    trans = ["C%d -> 0.1*C%d -> C%d" % (k, k, k + 1) for k in range(32)]
    m = si.mparse(trans, comps, u0={c: [1] for c in comps}, tspan=[1.0, 2.0])
    with pytest.raises(ValueError, match="at most 32 compartments"):
        plugins.device_source(m.plugin.parsed, "wide")
    # 32 compartments and 31 transitions are accepted.
    m = si.mparse(trans[:31], comps[:32], u0={c: [1] for c in comps[:32]},
                  tspan=[1.0, 2.0])
    assert "NC = 32, NT = 31" in plugins.device_source(m.plugin.parsed, "wide")

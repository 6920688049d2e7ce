"""Model parser: the "X -> propensity -> Y" language of the reference's
mparse, producing the stoichiometry S, the dependency graph G, the C code
of the transition-rate and post-time-step functions, and a gfx950 device
plugin compiled from the same function bodies.

Restates R/mparse.R (tokenize :27-85, rewrite_tokens :91-122,
propensity dependencies and variables :124-175, G labels :177-192,
parse_compartments :195-225, parse_propensity :227-254, variables
:274-356, topological_sort :372-399, S and G :475-491, mparse :647-724)
and R/C-generator.R:22-409 (the emitted C text, line for line).
"""
import re

import numpy as np

OPERATORS = ("<<=", ">>=", "!=", "%=", "##", "&&", "&=", "*=", "++", "+=",
             "--", "-=", "->", "/=", "<<", "<=", "==", ">=", ">>", "^=", "|=",
             "||", "!", "~", "%", "&", "(", ")", "*", "+", ",", "-", "/", ":",
             ";", "<", "=", ">", "?", "[", "]", "^", "{", "|", "}", "#")


class MparseError(ValueError):
    pass


def tokenize(code):
    """Split code into identifiers and operator/punctuator tokens: each
    operator in turn splits every non-operator piece (R/mparse.R:27-85)."""
    pieces = [(code, False)]
    for op in OPERATORS:
        out = []
        for tok, is_op in pieces:
            if is_op or op not in tok:
                out.append((tok, is_op))
                continue
            parts = []
            start = 0
            while True:
                k = tok.find(op, start)
                if k < 0:
                    break
                if k > start:
                    parts.append(tok[start:k])
                parts.append(op)
                start = k + len(op)
            if start < len(tok):
                parts.append(tok[start:])
            parts = [p.strip() for p in parts]
            out.extend((p, p == op) for p in parts if len(p) > 0)
        pieces = out
    return [t for t, _ in pieces]


def remove_spaces(x):
    return x.replace(" ", "")


class Names(list):
    """Variable names with their enumeration values (R attributes
    'value' and 'n_values')."""

    def __init__(self, names=(), values=None, n_values=None):
        super().__init__(names)
        self.values = list(values) if values is not None else None
        self.n_values = n_values if n_values is not None else len(self)


def variable_names(x, is_vector_ok, what):
    """Names of gdata/ldata/v0 (R/mparse.R:429-473).  x is None, a dict
    (named vector or data.frame columns), or a (names, values) pair."""
    if x is None:
        return None
    if isinstance(x, dict):
        lbl = [str(k) for k in x.keys()]
    elif isinstance(x, (list, tuple)) and len(x) == 2 and isinstance(x[0], (list, tuple)):
        lbl = ["" if k is None else str(k) for k in x[0]]
    else:
        raise MparseError("'%s' must either be a 'data.frame' or a 'matrix'." % what)
    value = list(range(len(lbl)))
    keep = [i for i, l in enumerate(lbl) if len(l) > 0]
    out = Names([lbl[i] for i in keep], [value[i] for i in keep], len(value))
    if len(set(out)) != len(out):
        raise MparseError("'%s' must have non-duplicated parameter names." % what)
    return out


def substitute_tokens(tokens, pattern, replacement, use_enum):
    if not pattern:
        return tokens
    out = []
    for t in tokens:
        if t in pattern:
            i = pattern.index(t)
            lbl = pattern[i].upper() if use_enum else str(i)
            out.append("%s[%s]" % (replacement, lbl))
        else:
            out.append(t)
    return out


def rewrite_tokens(tokens, compartments, ldata_names, gdata_names, v0_names,
                   use_enum):
    tokens = substitute_tokens(tokens, compartments, "u", use_enum)
    tokens = substitute_tokens(tokens, ldata_names, "ldata", use_enum)
    tokens = substitute_tokens(tokens, gdata_names, "gdata", use_enum)
    tokens = substitute_tokens(tokens, v0_names, "v", use_enum)
    return tokens


def _unique(seq):
    seen, out = set(), []
    for x in seq:
        if x not in seen:
            seen.add(x)
            out.append(x)
    return out


def propensity_dependencies(tokens, vars_, variables, compartments):
    depends = [0] * len(compartments)
    for t in tokens:
        if t in compartments:
            depends[compartments.index(t)] = 1
    names = list(variables.keys())
    todo = _unique(names.index(v) for v in vars_ if v in names)
    while todo:
        var = variables[names[todo[0]]]
        for c in var["compartments"]:
            if c in compartments:
                depends[compartments.index(c)] = 1
        todo = todo + [names.index(d) for d in var["depends"] if d in names]
        todo = _unique(todo[1:])
    return depends


def propensity_variables(tokens, variables):
    names = list(variables.keys())
    return _unique(t for t in tokens if t in names)


def rewrite_propensity(propensity, variables, compartments, ldata_names,
                       gdata_names, v0_names, use_enum):
    tokens = tokenize(propensity)
    G_rowname = "".join(tokens)
    vars_ = propensity_variables(tokens, variables)
    depends = propensity_dependencies(tokens, vars_, variables, compartments)
    tokens = rewrite_tokens(tokens, compartments, ldata_names, gdata_names,
                            v0_names, use_enum)
    return {"code": "".join(tokens), "depends": depends,
            "G_rowname": G_rowname, "variables": vars_}


def G_label(counts):
    """'@' for nothing, else 'n*X + Y' (R/mparse.R:178-188)."""
    if not counts:
        return "@"
    return " + ".join(("%d*" % n if n > 1 else "") + name for name, n in counts)


def parse_compartments(x, compartments):
    parts = [remove_spaces(p) for p in x.split("+")]
    expanded = []
    for p in parts:
        m = re.match(r"^[0-9]+[*]", p)
        if m:
            n = int(p[:m.end() - 1])
            expanded.extend([p[m.end():]] * n)
        else:
            expanded.append(p)
    if "@" in expanded and len(expanded) > 1:
        raise MparseError("Invalid usage of the empty set '@'.")
    expanded = [p for p in expanded if p != "@"]
    counts = [0] * len(compartments)
    for p in expanded:
        if p not in compartments:
            raise MparseError("Unknown compartment: '%s'." % p)
        counts[compartments.index(p)] += 1
    return counts


def parse_propensity(x, variables, compartments, ldata_names, gdata_names,
                     v0_names, use_enum):
    propensity = "->".join(remove_spaces(p) for p in x[1:-1])
    frm = parse_compartments(x[0], compartments)
    dst = parse_compartments(x[-1], compartments)
    S = [d - f for d, f in zip(dst, frm)]
    prop = rewrite_propensity(propensity, variables, compartments, ldata_names,
                              gdata_names, v0_names, use_enum)
    lf = G_label([(c, n) for c, n in zip(compartments, frm) if n > 0])
    ld = G_label([(c, n) for c, n in zip(compartments, dst) if n > 0])
    return {"code": prop["code"], "depends": prop["depends"], "S": S,
            "G_rowname": "%s -> %s -> %s" % (lf, prop["G_rowname"], ld),
            "variables": prop["variables"]}


PATTERN_VARIABLE = re.compile(
    r"^[ \t\n\r\f\v]*([(]double[)]|[(]int[)])?[ \t\n\r\f\v]*"
    r"([a-zA-Z_][a-zA-Z_0-9]*)[ \t\n\r\f\v]*<-")


def is_variable(transition):
    return PATTERN_VARIABLE.search(transition) is not None


def parse_variable(x, compartments, ldata_names, gdata_names, v0_names,
                   use_enum):
    m = PATTERN_VARIABLE.search(x)
    if m is None:
        raise MparseError("Invalid variable: '%s'." % x)
    vtype = m.group(1)[1:-1] if m.group(1) else "double"
    name = m.group(2)
    taken = list(compartments) + list(gdata_names or []) + \
        list(ldata_names or []) + list(v0_names or [])
    if name in taken:
        raise MparseError("Variable name already exists in 'u0', 'gdata', "
                          "'ldata' or 'v0'.")
    tokens = tokenize(remove_spaces(x[m.end():]))
    var_comps = _unique(t for t in tokens if t in compartments)
    tokens = rewrite_tokens(tokens, compartments, ldata_names, gdata_names,
                            v0_names, use_enum)
    return {"variable": name, "tokens": tokens, "type": vtype,
            "compartments": var_comps}


def topological_sort(depends, names):
    """Kahn's algorithm with a lexicographic start (R/mparse.R:372-399).
    depends[i][j] = 1 when variable names[j] uses names[i]."""
    order = sorted(names)
    idx = {n: names.index(n) for n in order}
    M = {c: {r: depends[idx[r]][idx[c]] for r in order} for c in order}
    result = []
    j = [c for c in order if sum(M[c].values()) == 0]
    if not j:
        raise MparseError("Invalid dependencies between variables.")
    cols = list(order)
    while j:
        first = j[0]
        result.append(first)
        cols.remove(first)
        for c in cols:
            M[c][first] = 0
        k = [c for c in cols if sum(M[c][r] for r in order) == 0]
        j = j + [c for c in k if c not in j]
        j = j[1:]
    if cols:
        raise MparseError("Invalid dependencies between variables.")
    return result


def parse_variables(variables, compartments, ldata_names, gdata_names,
                    v0_names, use_enum):
    if not variables:
        return {}
    parsed = [parse_variable(x, compartments, ldata_names, gdata_names,
                             v0_names, use_enum) for x in variables]
    names = [p["variable"] for p in parsed]
    if len(set(names)) != len(names):
        raise MparseError("Variables must have non-duplicated names.")
    n = len(names)
    dep = [[0] * n for _ in range(n)]
    for j, p in enumerate(parsed):
        for t in p["tokens"]:
            if t in names:
                dep[names.index(t)][j] = 1
    order = topological_sort(dep, names)
    out = {}
    for name in order:
        p = parsed[names.index(name)]
        j = names.index(name)
        p["depends"] = [o for o in order if dep[names.index(o)][j] > 0]
        p["code"] = "".join(p["tokens"])
        del p["tokens"]
        out[name] = p
    return out


def parse_transitions(transitions, compartments, ldata_names, gdata_names,
                      v0_names, use_enum):
    var_lines = [t for t in transitions if is_variable(t)]
    prop_lines = [t for t in transitions if not is_variable(t)]
    variables = parse_variables(var_lines, compartments, ldata_names,
                                gdata_names, v0_names, use_enum)
    props = []
    for t in prop_lines:
        x = t.split("->")
        if len(x) < 3:
            raise MparseError("Invalid transition: '%s'." % "->".join(x))
        props.append(parse_propensity(x, variables, compartments, ldata_names,
                                      gdata_names, v0_names, use_enum))
    return props, variables


# ------------------------------------------------------- C generator ---

def C_heading(version="10.1.0"):
    return ["/* Generated by SimInf (v%s) */" % version, ""]


def C_include():
    return ["#include <R_ext/Rdynload.h>", "#include \"SimInf.h\"", ""]


def C_define():
    return [
        "/**",
        " * Make sure the necessary macros are defined so that the",
        " * compiler can replace them when compiling the model.",
        " * 'SIMINF_MODEL_RUN' defines the function name of the function",
        " * that will be called from R to run a trajectory of the model.",
        " * 'SIMINF_R_INIT' is the name of the function that R will call",
        " * when this model is loaded into R. 'SIMINF_FORCE_SYMBOLS'",
        " * defines whether R allows the entry point for the run function",
        " * to be searched for as a character string.",
        " * If this file is compiled from SimInf (when calling run), the",
        " * macros are defined by SimInf before calling 'R CMD SHLIB'.",
        " * If this file is compiled as part of a package, then the",
        " * definitions are set in the variable 'PKG_CPPFLAGS' in",
        " * 'src/Makevars' and 'src/Makevars.in'.",
        " */",
        "#if !defined(SIMINF_MODEL_RUN)",
        "#  error Definition for 'SIMINF_MODEL_RUN' is missing.",
        "#endif",
        "#if !defined(SIMINF_R_INIT)",
        "#  error Definition for 'SIMINF_R_INIT' is missing.",
        "#endif",
        "#if !defined(SIMINF_FORCE_SYMBOLS)",
        "#  error Definition for 'SIMINF_FORCE_SYMBOLS' is missing.",
        "#endif",
        ""]


def C_enumeration_constants(target, labels):
    if not labels:
        return []
    lines = ["/* Enumeration constants for indicies in the '%s' vector. */"
             % target, "enum {"]
    for i, lbl in enumerate(labels):
        s = lbl.upper()
        if labels.values is not None:
            s = "%s = %d" % (s, labels.values[i])
        if i < len(labels) - 1:
            s += ","
        lines.append("    " + s)
    return lines + ["};", ""]


def C_enum(compartments, ldata_names, gdata_names, v0_names, use_enum):
    if not use_enum:
        return []
    comps = None
    if compartments:
        comps = Names(list(compartments) + ["N_COMPARTMENTS_U"],
                      list(compartments.values) + [compartments.n_values],
                      compartments.n_values + 1)
    v0 = None
    if v0_names:
        v0 = Names(list(v0_names) + ["N_COMPARTMENTS_V"],
                   list(v0_names.values) + [v0_names.n_values],
                   v0_names.n_values + 1)
    return (C_enumeration_constants("u", comps) +
            C_enumeration_constants("v", v0) +
            C_enumeration_constants("ldata", ldata_names) +
            C_enumeration_constants("gdata", gdata_names))


def C_pre_code(pre_code):
    if pre_code is None:
        return []
    lines = []
    for chunk in ([pre_code] if isinstance(pre_code, str) else pre_code):
        lines.extend(chunk.split("\n"))
    return ["/* User defined code. */"] + lines + [""]


def C_variables(prop, variables):
    names = list(variables.keys())
    j = [names.index(v) for v in prop["variables"] if v in names]
    i = []
    while j:
        if j[0] not in i:
            i = sorted(i + [j[0]])
            j = j + [names.index(d) for d in variables[names[j[0]]]["depends"]]
        j = j[1:]
    lines = ["    const %s %s = %s;" % (variables[names[k]]["type"], names[k],
                                        variables[names[k]]["code"]) for k in i]
    return lines + [""] if lines else lines


def C_trFun(props, variables):
    params = ["    const int *u,", "    const double *v,", "    const double *ldata,",
              "    const double *gdata,", "    double t)"]
    lines = []
    for i, p in enumerate(props, start=1):
        lines += ["/**",
                  " * @param u The compartment state vector in the node.",
                  " * @param v The continuous state vector in the node.",
                  " * @param ldata The local data vector in the node.",
                  " * @param gdata The global data vector.",
                  " * @param t Current time.",
                  " * @return propensity.",
                  " */",
                  "static double trFun%d(" % i] + params + ["{"]
        lines += C_variables(p, variables)
        lines += ["    return %s;" % p["code"], "}", ""]
    return lines


def _pts_lines(pts_fun):
    if pts_fun is None:
        pts_fun = "    return 0;"
    lines = []
    for chunk in ([pts_fun] if isinstance(pts_fun, str) else pts_fun):
        lines.extend(chunk.split("\n"))
    return lines


def C_ptsFun(pts_fun):
    return ["/**",
            " * Post time step function.",
            " *",
            " * @param v_new If a continuous state vector is used by a model,",
            " *        this is the new continuous state vector in the node after",
            " *        the post time step.",
            " * @param u The compartment state vector in the node.",
            " * @param v The current continuous state vector in the node.",
            " * @param ldata The local data vector in the node.",
            " * @param gdata The global data vector that is common to all nodes.",
            " * @param node The node index. Note the node index is zero-based,",
            " *        i.e., the first node is 0.",
            " * @param t Current time in the simulation.",
            " * @return error code (<0), or 1 if node needs to update the",
            " *         transition rates, or 0 when it doesn't need to update",
            " *         the transition rates.",
            " */",
            "static int ptsFun(",
            "    double *v_new,",
            "    const int *u,",
            "    const double *v,",
            "    const double *ldata,",
            "    const double *gdata,",
            "    int node,",
            "    double t)",
            "{"] + _pts_lines(pts_fun) + ["}", ""]


def C_run(n_tr):
    return ["/**",
            " * Run a trajectory of the model.",
            " *",
            " * @param model The model.",
            " * @param options Options for running the numerical solver.",
            " * @return A model with a trajectory attached to it.",
            " */",
            "static SEXP SIMINF_MODEL_RUN(SEXP model, SEXP options)",
            "{",
            "    static SEXP(*SimInf_run)(SEXP, SEXP, TRFun*, PTSFun) = NULL;",
            "    TRFun tr_fun[] = {%s};" % ", ".join("&trFun%d" % i
                                                  for i in range(1, n_tr + 1)),
            "",
            "    if (!SimInf_run) {",
            "        SimInf_run = (SEXP(*)(SEXP, SEXP, TRFun*, PTSFun))",
            "            R_GetCCallable(\"SimInf\", \"SimInf_run\");",
            "",
            "        if (!SimInf_run) {",
            "            Rf_error(\"Cannot find function 'SimInf_run'.\");",
            "        }",
            "    }",
            "",
            "    return SimInf_run(model, options, tr_fun, &ptsFun);",
            "}",
            ""]


def C_calldef():
    return ["/**",
            " * A NULL-terminated array of routines to register for the .Call",
            " * interface, see section '5.4 Registering native routines' in",
            " * the 'Writing R Extensions' manual.",
            " */",
            "static const R_CallMethodDef callMethods[] =",
            "{",
            "    SIMINF_CALLDEF(SIMINF_MODEL_RUN, 2),",
            "    {NULL, NULL, 0}",
            "};",
            ""]


def C_R_init():
    return ["/**",
            " * This routine will be invoked when R loads the shared object/DLL,",
            " * see section '5.4 Registering native routines' in the",
            " * 'Writing R Extensions' manual.",
            " */",
            "void SIMINF_R_INIT(DllInfo *info)",
            "{",
            "    R_registerRoutines(info, NULL, callMethods, NULL, NULL);",
            "    R_useDynamicSymbols(info, FALSE);",
            "    R_forceSymbols(info, SIMINF_FORCE_SYMBOLS);",
            "}"]


def C_code_mparse(props, variables, pts_fun, compartments, ldata_names,
                  gdata_names, v0_names, use_enum, pre_code):
    return (C_heading() + C_include() + C_define() +
            C_enum(compartments, ldata_names, gdata_names, v0_names, use_enum) +
            C_pre_code(pre_code) + C_trFun(props, variables) +
            C_ptsFun(pts_fun) + C_run(len(props)) + C_calldef() + C_R_init())


class ParsedModel:
    """What mparse derives from the transitions: S (Nc x Nt), G
    (Nt x Nt) with row labels, and the C code lines."""

    def __init__(self, transitions, compartments, ldata=None, gdata=None,
                 v0=None, pts_fun=None, use_enum=False, pre_code=None,
                 single_node=True):
        if isinstance(transitions, str):
            transitions = [transitions]
        if not transitions or any(len(t) == 0 for t in transitions):
            raise MparseError("'transitions' must be specified in a character vector.")
        comps = Names(list(compartments), range(len(compartments)))
        ldn = variable_names(ldata, single_node, "ldata")
        gdn = variable_names(gdata, True, "gdata")
        v0n = variable_names(v0, single_node, "v0")
        everything = list(comps) + list(gdn or []) + list(ldn or []) + list(v0n or [])
        if len(set(everything)) != len(everything):
            raise MparseError("'u0', 'gdata', 'ldata' and 'v0' have names in common.")
        if {"N_COMPARTMENTS_U", "N_COMPARTMENTS_V"} & set(everything):
            raise MparseError("Invalid compartment or variable name.")
        props, variables = parse_transitions(list(transitions), comps, ldn, gdn,
                                             v0n, use_enum)
        self.props = props
        self.variables = variables
        self.compartments = list(compartments)
        self.S = np.array([p["S"] for p in props], dtype=float).T
        depends = np.array([p["depends"] for p in props], dtype=float)
        self.G = ((depends @ np.abs(self.S)) > 0).astype(float)
        self.G_rownames = [p["G_rowname"] for p in props]
        self.C_code = C_code_mparse(props, variables, pts_fun, comps, ldn, gdn,
                                    v0n, use_enum, pre_code)
        self.pts_fun = pts_fun
        self.pre_code = pre_code
        self.use_enum = use_enum
        self.compartments_names = comps
        self.ldata_names, self.gdata_names, self.v0_names = ldn, gdn, v0n
        self.Nld = ldn.n_values if ldn is not None else 0
        self.Ngdata = gdn.n_values if gdn is not None else 0
        self.Nd = v0n.n_values if v0n is not None else 0


def _named_rows(x, n_nodes):
    """(names, Nrows x Nn matrix) from a dict of per-node columns (the
    data.frame form), a dict of scalars (one node), or None."""
    if x is None:
        return None, np.zeros((0, n_nodes))
    names = list(x.keys())
    rows = [np.broadcast_to(np.atleast_1d(np.asarray(x[k], dtype=float)), (n_nodes,))
            for k in names]
    return names, np.vstack(rows) if rows else np.zeros((0, n_nodes))


def mparse(transitions, compartments, ldata=None, gdata=None, u0=None, v0=None,
           tspan=None, events=None, E=None, N=None, pts_fun=None,
           use_enum=False, pre_code=None, replicates=1):
    """Parse a model description into a SimInfModel whose plugin is
    generated from the transitions (R/mparse.R:647-724).

    u0, ldata and v0 map names to per-node values (data.frame columns);
    gdata maps names to values (a named vector; None keys are unnamed).
    """
    from .model import SimInfModel
    from .plugins import GeneratedPlugin
    if u0 is None:
        raise MparseError("Missing columns in initial state.")
    missing = [c for c in compartments if c not in u0]
    if missing:
        raise MparseError("Missing columns in initial state.")
    cols = [np.atleast_1d(np.asarray(u0[c])) for c in compartments]
    n_nodes = max(len(c) for c in cols)
    u = np.vstack([np.broadcast_to(c, (n_nodes,)) for c in cols]).astype(np.int64)
    ld_names, ld = _named_rows(ldata, n_nodes)
    v_names, v = _named_rows(v0, n_nodes)
    if gdata is None:
        g_pairs = None
        gvec = []
    else:
        keys = list(gdata.keys())
        g_pairs = ([k if k is not None else "" for k in keys], None)
        gvec = [float(gdata[k]) for k in keys]
    parsed = ParsedModel(transitions, compartments,
                         ldata={k: None for k in ld_names} if ld_names else None,
                         gdata=(g_pairs[0], gvec) if g_pairs else None,
                         v0={k: None for k in v_names} if v_names else None,
                         pts_fun=pts_fun, use_enum=use_enum, pre_code=pre_code,
                         single_node=n_nodes == 1)
    if replicates > 1:
        u = np.tile(u, (1, replicates))
        v = np.tile(v, (1, replicates))
    model = SimInfModel(GeneratedPlugin(parsed), parsed.G, parsed.S, tspan, u,
                        ldata=ld, gdata=gvec, v0=v if v.shape[0] else None,
                        E=E if E is not None else np.zeros((len(compartments), 0)),
                        N=N, events=events, replicates=replicates,
                        compartments=compartments, v0_names=v_names,
                        ldata_names=ld_names,
                        gdata_names=[k if k is not None else "" for k in gdata]
                        if gdata else None)
    model.C_code = parsed.C_code
    model.G_rownames = parsed.G_rownames
    return model

"""Extract golden vectors from the reference's own R tests into JSON
fixtures (data only).  Run in the build container, where the reference
checkout is mounted at /root/reference; the GPU box only reads the JSON.

  seir_enter_events.json  tests/SEIR.R:231-295 (U after one enter event
                          per compartment, node and day; zero rates)
  seir_zero_rates.json    tests/SEIR.R:134-171 (S/I/R with zero rates)
  phi_sise.json           tests/phi.R:34-105 (phi of an SISe node with
                          upsilon = gamma = alpha = 0)
  prevalence.json         tests/prevalence.R:31-95 (a hand-made U with
                          conditional prevalences) and tests/SEIR.R:327-410
                          (levels 1-3 of the enter-events trajectory)
"""
import json
import os
import re

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def _numbers(text):
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:e-?\d+)?", text)]


def _vector_after(src, start, name):
    """Numbers of 'name = c(...)', 'name <- c(...)' or 'name(c(...)' found
    after start."""
    m = re.compile(r"\b%s\s*(?:=|<-|\()\s*c\(" % re.escape(name)).search(src, start)
    depth, i = 1, m.end()
    while depth:
        depth += {"(": 1, ")": -1}.get(src[i], 0)
        i += 1
    return _numbers(src[m.end():i - 1]), i


def _r_strings(src, start):
    """The string literals of the c(...) that opens at src[start:]."""
    i = src.index("c(", start) + 2
    out, depth = [], 1
    while depth:
        ch = src[i]
        if ch == '"':
            j, buf = i + 1, []
            while src[j] != '"':
                if src[j] == "\\":
                    buf.append({"n": "\n", "t": "\t"}.get(src[j + 1], src[j + 1]))
                    j += 2
                else:
                    buf.append(src[j])
                    j += 1
            out.append("".join(buf))
            i = j + 1
            continue
        depth += {"(": 1, ")": -1}.get(ch, 0)
        i += 1
    return out


def main():
    seir = open(os.path.join(REF, "SEIR.R")).read()
    start = seir.index("U_expected <- data.frame(")
    out = {"source": "tests/SEIR.R:256-292"}
    pos = start
    for col in ("node", "time", "S", "E", "I", "R"):
        vals, pos = _vector_after(seir, pos, col)
        out[col] = [int(v) for v in vals]
    json.dump(out, open(os.path.join(HERE, "seir_enter_events.json"), "w"))

    start = seir.index("S_expected <- structure(")
    zero = {"source": "tests/SEIR.R:143-168"}
    for name, key in (("S_expected", "S"), ("I_expected", "I"), ("R_expected", "R")):
        i = seir.index(name + " <- structure(", start)
        vals, _ = _vector_after(seir, i, "structure")
        zero[key] = [int(v) for v in vals]
    json.dump(zero, open(os.path.join(HERE, "seir_zero_rates.json"), "w"))

    mp = open(os.path.join(REF, "mparse.R")).read()
    blocks = [m.start() for m in re.finditer(r"\nC_code <- c\(", mp)]
    json.dump({"source": "tests/mparse.R C_code blocks at lines %s" % [
                   mp.count("\n", 0, b) + 2 for b in blocks],
               "blocks": [_r_strings(mp, b) for b in blocks]},
              open(os.path.join(HERE, "mparse_c_code.json"), "w"), indent=0)

    phi = open(os.path.join(REF, "phi.R")).read()
    vals, _ = _vector_after(phi, 0, "phi_exp")
    json.dump({"source": "tests/phi.R:34-83", "tol": 1e-8, "phi": vals},
              open(os.path.join(HERE, "phi_sise.json"), "w"))

    json.dump(_prevalence_fixtures(seir), open(os.path.join(HERE, "prevalence.json"), "w"),
              indent=0)


def _fractions(text):
    """Values of a c(a / b, ...) of integer fractions."""
    out = []
    for item in text.split(","):
        num, den = item.split("/")
        out.append(int(num) / int(den))
    return out


def _prevalence_fixtures(seir):
    src = open(os.path.join(REF, "prevalence.R")).read()
    i = src.index("model@U <- matrix(c(")
    U, _ = _vector_after(src, i, "matrix")
    out = {"source": "tests/prevalence.R:31-95 and tests/SEIR.R:327-410",
           "hand": {"U": [int(v) for v in U], "nrow": 9, "ncol": 5, "byrow": True,
                    "compartments": ["S", "I", "R"], "cases": []}}
    call = re.compile(r"p <- prevalence\(model, (.+?)\)\$prevalence\n(.+?)(?=\n\n|\np <-|\n?\Z)",
                      re.S)
    for m in call.finditer(src):
        args, checks = m.group(1), m.group(2)
        formula, index = args, None
        if ", i = " in args:
            formula, idx = args.split(", i = ")
            index = [int(idx)]
        expected = [None] * 5
        for c in re.finditer(r"abs\(p(\[(\d):(\d)\])? - c\(([^)]*)\)\)", checks):
            lo = int(c.group(2)) - 1 if c.group(1) else 0
            vals = _fractions(c.group(4))
            expected[lo:lo + len(vals)] = vals
        out["hand"]["cases"].append({"formula": formula, "index": index,
                                     "expected": expected})
    src.index("all(is.nan(prevalence(model, I ~ . | R == 5)$prevalence))")
    out["hand"]["cases"].append({"formula": "I ~ . | R == 5", "index": None,
                                 "expected": [None] * 5})

    start = seir.index("## Check prevalence")
    lvl1, pos = _vector_after(seir, start, "prevalence")
    lvl2, pos = _vector_after(seir, pos, "prevalence")
    node3, pos = _vector_after(seir, pos, "node")
    time3, pos = _vector_after(seir, pos, "time")
    prev3, pos = _vector_after(seir, pos, "prevalence")
    nodes_sub, pos = _vector_after(seir, pos, "node")
    time_sub, pos = _vector_after(seir, pos, "time")
    prev_sub, pos = _vector_after(seir, pos, "prevalence")
    out["seir"] = {"tol": 1e-8, "level1": lvl1, "level2": lvl2,
                   "level3": {"node": [int(v) for v in node3], "time": time3,
                              "prevalence": prev3},
                   "level3_index_2_3": {"node": [int(v) for v in nodes_sub],
                                        "time": time_sub, "prevalence": prev_sub}}
    return out


if __name__ == "__main__":
    main()

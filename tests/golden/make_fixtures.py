"""Extract golden vectors from the reference's own R tests into JSON
fixtures (data only).  Run in the build container, where the reference
checkout is mounted at /root/reference; the GPU box only reads the JSON.

  seir_enter_events.json  tests/SEIR.R:231-295 (U after one enter event
                          per compartment, node and day; zero rates)
  seir_zero_rates.json    tests/SEIR.R:134-171 (S/I/R with zero rates)
  phi_sise.json           tests/phi.R:34-105 (phi of an SISe node with
                          upsilon = gamma = alpha = 0)
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


if __name__ == "__main__":
    main()

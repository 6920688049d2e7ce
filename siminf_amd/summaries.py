"""Specifications of the on-device summaries: prevalence formulas and
compartment selections, checked with the reference's messages and turned
into the masks and the condition program of `SimInf_hip_plan_prevalence`
/ `SimInf_hip_plan_trajectory`.

Mirrors (what, not how):
  parse_formula / parse_formula_item / replace_dot  R/match_compartments.R:24-93
  prevalence argument checks                        R/prevalence.R:233-244
  check_node_index_argument                         R/check_arguments.R:202-214
  evaluate_condition                                R/prevalence.R:40-68

A formula is given as text, ``"cases ~ population | condition"``, with
``.`` for every compartment.  The condition is an R expression over
compartment names: numbers, TRUE/FALSE, ``+ - * /``, unary ``- + !``,
comparisons ``== != < <= > >=`` and ``& && | ||``, with R's precedence
(arithmetic, then comparison, then ``!``, then ``&``, then ``|``).  It must
give a logical for every node and time step, so it must read at least one
compartment.  It compiles to a postfix program the device evaluates per
node and column in double precision.
"""
import re

import numpy as np

OP_U, OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_NEG = range(7)
OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE, OP_AND, OP_OR, OP_NOT = range(7, 16)
COND_MAX = 64
COND_STACK = 16

_BINARY = {"+": OP_ADD, "-": OP_SUB, "*": OP_MUL, "/": OP_DIV,
           "==": OP_EQ, "!=": OP_NE, "<": OP_LT, "<=": OP_LE, ">": OP_GT,
           ">=": OP_GE, "&": OP_AND, "&&": OP_AND, "|": OP_OR, "||": OP_OR}
_COMPARE = ("==", "!=", "<", "<=", ">", ">=")
_CONDITION_ERROR = ("The condition must be either 'TRUE' or 'FALSE' for every "
                    "node and time step.")


class SummaryError(ValueError):
    """A summary specification the reference rejects (same message)."""


def parse_formula_item(x):
    """'a + b + ...' -> unique trimmed names (R/match_compartments.R:24-30)."""
    out = []
    for item in x.split("+"):
        item = item.strip()
        if item and item not in out:
            out.append(item)
    if not out:
        raise SummaryError("No compartments in formula specification.")
    return out


def _replace_dot(items, compartments):
    out = []
    for x in items:
        for y in (compartments if x == "." else [x]):
            if y not in out:
                out.append(y)
    return out


def parse_formula(formula, compartments):
    """(lhs names, rhs names, condition text or None) of 'lhs ~ rhs | cond';
    lhs is None for a one-sided formula."""
    if formula.count("~") != 1:
        raise SummaryError("Invalid formula specification of 'compartments'.")
    lhs_text, rhs = formula.split("~")
    lhs = parse_formula_item(lhs_text) if lhs_text.strip() else None
    condition = None
    if "|" in rhs and rhs.index("|") > 0:
        rhs, condition = rhs.split("|", 1)
        condition = condition.strip()
        if not condition:
            raise SummaryError("Invalid formula specification of 'condition'.")
    rhs = _replace_dot(parse_formula_item(rhs), compartments)
    if lhs is not None:
        lhs = _replace_dot(lhs, compartments)
    missing = [x for x in (lhs or []) + rhs if x not in compartments]
    if missing:
        raise SummaryError("Non-existing compartment(s) in model: %s." %
                           ", ".join("'%s'" % x for x in dict.fromkeys(missing)))
    return lhs, rhs, condition


def mask(names, compartments):
    m = 0
    for x in names:
        m |= 1 << compartments.index(x)
    return m


_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?L?|\.\d+(?:[eE][-+]?\d+)?)"
                    r"|([A-Za-z.][A-Za-z0-9._]*)|(&&|\|\||==|!=|<=|>=|[-+*/<>!&|()]))")


def _tokens(text):
    pos, out = 0, []
    text = text.rstrip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise SummaryError("Invalid formula specification of 'condition'.")
        pos = m.end()
        if m.group(1) is not None:
            out.append(("num", float(m.group(1).rstrip("L"))))
        elif m.group(2) is not None:
            out.append(("name", m.group(2)))
        else:
            out.append(("op", m.group(3)))
    return out


class _Compiler:
    """Recursive descent over R's precedence, emitting postfix code and
    tracking whether each subexpression is logical."""

    def __init__(self, tokens, compartments):
        self.t = tokens
        self.i = 0
        self.comp = compartments
        self.ops, self.args = [], []
        self.reads = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, *ops):
        kind, val = self.peek()
        if kind == "op" and val in ops:
            self.i += 1
            return val
        return None

    def emit(self, op, arg=0.0):
        self.ops.append(op)
        self.args.append(float(arg))

    def parse(self):
        logical = self.or_()
        if self.i != len(self.t):
            raise SummaryError("Invalid formula specification of 'condition'.")
        return logical

    def or_(self):
        logical = self.and_()
        while self.take("|", "||"):
            self.and_()
            self.emit(OP_OR)
            logical = True
        return logical

    def and_(self):
        logical = self.not_()
        while self.take("&", "&&"):
            self.not_()
            self.emit(OP_AND)
            logical = True
        return logical

    def not_(self):
        if self.take("!"):
            self.not_()
            self.emit(OP_NOT)
            return True
        return self.compare()

    def compare(self):
        logical = self.additive()
        op = self.take(*_COMPARE)
        if op:
            self.additive()
            self.emit(_BINARY[op])
            if self.take(*_COMPARE):
                raise SummaryError("Invalid formula specification of 'condition'.")
            return True
        return logical

    def additive(self):
        logical = self.multiplicative()
        while True:
            op = self.take("+", "-")
            if not op:
                return logical
            self.multiplicative()
            self.emit(_BINARY[op])
            logical = False

    def multiplicative(self):
        logical = self.unary()
        while True:
            op = self.take("*", "/")
            if not op:
                return logical
            self.unary()
            self.emit(_BINARY[op])
            logical = False

    def unary(self):
        op = self.take("-", "+")
        if op:
            self.unary()
            if op == "-":
                self.emit(OP_NEG)
            return False
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        self.i += 1
        if kind == "num":
            self.emit(OP_CONST, val)
            return False
        if kind == "name":
            if val in ("TRUE", "T", "FALSE", "F"):
                self.emit(OP_CONST, 1.0 if val in ("TRUE", "T") else 0.0)
                return True
            if val not in self.comp:
                raise SummaryError("Non-existing compartment(s) in model: '%s'." % val)
            self.emit(OP_U, self.comp.index(val))
            self.reads += 1
            return False
        if kind == "op" and val == "(":
            logical = self.or_()
            if not self.take(")"):
                raise SummaryError("Invalid formula specification of 'condition'.")
            return logical
        raise SummaryError("Invalid formula specification of 'condition'.")


def stack_depth(ops):
    depth = peak = 0
    for op in ops:
        if op in (OP_U, OP_CONST):
            depth += 1
        elif op not in (OP_NEG, OP_NOT):
            depth -= 1
        peak = max(peak, depth)
    return peak


def compile_condition(text, compartments):
    """(ops int32, args float64) of a condition; raises the reference's
    message for a condition that is not a logical for every node and time
    step."""
    c = _Compiler(_tokens(text), list(compartments))
    logical = c.parse()
    if not logical or c.reads == 0:
        raise SummaryError(_CONDITION_ERROR)
    if len(c.ops) > COND_MAX or stack_depth(c.ops) > COND_STACK:
        raise SummaryError("The condition is too long for the device "
                           "(at most %d terms, nesting depth %d)." % (COND_MAX, COND_STACK))
    return np.asarray(c.ops, dtype=np.int32), np.asarray(c.args, dtype=np.float64)


def prevalence_spec(formula, compartments, level):
    """Checked (cases_mask, pop_mask, ops, args, level) of prevalence()."""
    lhs, rhs, condition = parse_formula(formula, compartments)
    if lhs is None:
        raise SummaryError("Invalid 'formula' specification.")
    if isinstance(level, bool) or not np.isscalar(level) or level != int(level) \
            or int(level) not in (1, 2, 3):
        raise SummaryError("'level' must be an integer with a value 1, 2 or 3.")
    ops = np.zeros(0, dtype=np.int32)
    args = np.zeros(0, dtype=np.float64)
    if condition is not None:
        ops, args = compile_condition(condition, compartments)
    return mask(lhs, compartments), mask(rhs, compartments), ops, args, int(level)


def node_index(index, n_nodes):
    """Zero-based sorted unique nodes of a one-based index, or None
    (check_node_index_argument, R/check_arguments.R:202-214)."""
    if index is None:
        return None
    idx = np.atleast_1d(np.asarray(index))
    if idx.size == 0 or idx.dtype.kind not in "iuf" or np.any(idx != np.round(idx)) \
            or idx.min() < 1 or idx.max() > n_nodes:
        raise SummaryError("The node index must be an integer > 0 and <= number of nodes.")
    return np.unique(idx.astype(np.int64)).astype(np.int32) - 1


def selected_names(compartments_arg):
    """The distinct, stripped names of a compartments argument (None for
    all)."""
    if compartments_arg is None:
        return None
    names = [compartments_arg] if isinstance(compartments_arg, str) else list(compartments_arg)
    return list(dict.fromkeys(n.strip() for n in names))


def trajectory_mask(compartments_arg, compartments):
    """Mask of the compartments to extract (all when None); the rows come
    out in model order."""
    if compartments_arg is None:
        return mask(compartments, compartments)
    names = [compartments_arg] if isinstance(compartments_arg, str) else list(compartments_arg)
    names = list(dict.fromkeys(n.strip() for n in names))
    missing = [x for x in names if x not in compartments]
    if missing:
        raise SummaryError("Non-existing compartment(s) in model: %s." %
                           ", ".join("'%s'" % x for x in missing))
    return mask(names, compartments)

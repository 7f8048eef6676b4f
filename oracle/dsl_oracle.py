"""CPU oracle, structured-product language -- TEST INFRASTRUCTURE ONLY (the
reference results the GPU structured-product engine is tested against; never imported by the
product).

Restates src/dsl of rosssaunders/openferric in plain Python:
  * tokenize (lexer.rs:80-380): indentation-driven INDENT / DEDENT, keywords,
    numbers with '_' separators, strings, // comments;
  * parse (parser.rs:158-700): product items, statements, precedence
    climbing (or < and < not < comparison < additive < multiplicative <
    unary minus < primary);
  * compile (compiler.rs:40-330): slot resolution (a new local slot per
    `let`, locals reset per schedule, state slots in declaration order),
    ScheduleFreq::generate_dates (ast.rs:64-80);
  * the evaluator's semantics (eval.rs:268-500, 1774-2010), walked over the
    tree instead of the reference's stack bytecode: observation -> snapshot
    step round(t n / T) capped at n, locals cleared per observation, pay
    adds amount * exp(-r t), redeem pays and ends the path, skip ends it,
    schedules run in order and an ended path stops, == / != against
    f64::EPSILON, division by zero is NaN, truthiness v != 0, worst_of /
    best_of of argument lists fold from the last argument, price(i) with
    `as usize` saturation and the reference's range error;
  * DslMonteCarloEngine::price_multi_asset (engine.rs:255-495): validation
    order and messages, AssetStepper (GBM with r - q or r_d - r_f,
    Schwartz one-factor, Vasicek), correlated normals L z summed from 0, the
    single xoshiro256++ stream path-major (path, step, asset), and the
    chunk close (sum, sum of squares, unbiased variance).

Paths are simulated with numpy from explicit normals [path][step][asset] or
from the engine's single stream (oracle.lsm_stream_normals); the evaluation
is a pure-Python loop, so keep the path counts small.

Pinned by tests/test_dsl_oracle.py: closed forms (zero coupon, forward,
Black-Scholes call and digital, the Vasicek and Schwartz means, a
two-asset worst-of digital against the bivariate normal), agreement with
the independently pinned autocallable oracle (or_autocall_price) on the
same normals for autocallables written in the language, and the
reference's own DSL tests (lexer, parser, compiler, evaluator and engine
unit tests, the documented examples).
"""
import math
import re

import numpy as np

EPS = 2.220446049250313e-16


class DslError(Exception):
    """DslError (src/dsl/error.rs:17-46); str() is the reference's text."""


class PricingError(Exception):
    def __init__(self, kind, msg):
        super().__init__(msg)
        self.kind = kind


# ---------------------------------------------------------------- lexer --

KEYWORDS = {w: w.capitalize() for w in (
    "product notional maturity underlyings state schedule from to let if then else pay "
    "redeem set skip and or not asset equity fx commodity rate true false monthly quarterly "
    "annual bool float").split()}
KEYWORDS["semi_annual"] = "SemiAnnual"


def _lex_line(content, offset, out):
    pos, n = 0, len(content)
    while pos < n:
        while pos < n and content[pos] == " ":
            pos += 1
        if pos >= n or content.startswith("//", pos):
            break
        start, ch = offset + pos, content[pos]
        if ch == '"':
            p, s = pos + 1, []
            while p < n and content[p] != '"':
                if content[p] == "\\" and p + 1 < n:
                    p += 1
                    s.append({"n": "\n", "t": "\t"}.get(content[p], content[p]))
                else:
                    s.append(content[p])
                p += 1
            if p >= n:
                raise DslError(f"lex error at {offset + pos}-{offset + p}: unterminated string literal")
            out.append(("StringLit", "".join(s), start, offset + p + 1))
            pos = p + 1
            continue
        if ch.isdigit() or (ch == "." and pos + 1 < n and content[pos + 1].isdigit()):
            m = re.match(r"[0-9_]*(\.[0-9_]*)?([eE][+-]?[0-9]*)?", content[pos:])
            text = m.group(0)
            num = text.replace("_", "")
            if re.search(r"[eE][+-]?$", num):   # exponent without digits
                raise DslError(f"lex error at {start}-{start + len(text)}: invalid number literal "
                               f"'{num.replace('E', 'e')}'")
            out.append(("Number", float(num), start, start + len(text)))
            pos += len(text)
            continue
        if ch.isalpha() or ch == "_":
            m = re.match(r"[A-Za-z0-9_]+", content[pos:])
            word = m.group(0)
            kind = KEYWORDS.get(word, "Ident")
            out.append((kind, word, start, start + len(word)))
            pos += len(word)
            continue
        two = content[pos:pos + 2]
        if two in ("==", "!=", "<=", ">="):
            out.append(({"==": "EqEq", "!=": "Ne", "<=": "Le", ">=": "Ge"}[two], two, start,
                        start + 2))
            pos += 2
            continue
        one = {"(": "LParen", ")": "RParen", ":": "Colon", ",": "Comma", "=": "Eq", "<": "Lt",
               ">": "Gt", "+": "Plus", "-": "Minus", "*": "Star", "/": "Slash"}.get(ch)
        if one is None:
            raise DslError(f"lex error at {start}-{start + 1}: unexpected character '{ch}'")
        out.append((one, ch, start, start + 1))
        pos += 1


def tokenize(src):
    toks, stack, pos = [], [0], 0
    for line in src.split("\n"):
        line_start = pos
        clean = line.rstrip("\r")
        indent = len(clean) - len(clean.lstrip(" "))
        content = clean[indent:].rstrip()
        pos += len(line) + 1
        if not content or content.startswith("//"):
            continue
        if indent > stack[-1]:
            stack.append(indent)
            toks.append(("Indent", None, line_start, line_start + indent))
        elif indent < stack[-1]:
            while stack[-1] > indent:
                stack.pop()
                toks.append(("Dedent", None, line_start, line_start + indent))
            if stack[-1] != indent:
                raise DslError(f"lex error at {line_start}-{line_start + indent}: inconsistent "
                               f"indentation: expected {stack[-1]} spaces, got {indent}")
        _lex_line(content, line_start + indent, toks)
    while len(stack) > 1:
        stack.pop()
        toks.append(("Dedent", None, pos, pos))
    return toks


# --------------------------------------------------------------- parser --

class _P:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def kind(self):
        return self.t[self.i][0] if self.i < len(self.t) else None

    def take(self, kind=None, what=None):
        if self.i >= len(self.t):
            end = self.t[-1][3] if self.t else 0
            raise DslError(f"parse error at {end}-{end}: expected {what or kind}, got end of input")
        tok = self.t[self.i]
        if kind is not None and tok[0] != kind:
            raise DslError(f"parse error at {tok[2]}-{tok[3]}: expected {what or kind}, got "
                           f"{tok[0]}")
        self.i += 1
        return tok

    def block_end(self):
        return self.kind() in (None, "Dedent")


def _expr(p):
    return _binary(p, 0)


_LEVELS = [("Or",), ("And",)]
_CMP = {"EqEq": "==", "Ne": "!=", "Lt": "<", "Le": "<=", "Gt": ">", "Ge": ">="}


def _binary(p, level):
    if level == 0:
        left = _binary(p, 1)
        while p.kind() == "Or":
            p.take()
            left = ("bin", "or", left, _binary(p, 1))
        return left
    if level == 1:
        left = _not(p)
        while p.kind() == "And":
            p.take()
            left = ("bin", "and", left, _not(p))
        return left


def _not(p):
    if p.kind() == "Not":
        p.take()
        return ("un", "not", _not(p))
    left = _additive(p)
    if p.kind() in _CMP:
        op = _CMP[p.take()[0]]
        return ("bin", op, left, _additive(p))
    return left


def _additive(p):
    left = _mult(p)
    while p.kind() in ("Plus", "Minus"):
        op = "+" if p.take()[0] == "Plus" else "-"
        left = ("bin", op, left, _mult(p))
    return left


def _mult(p):
    left = _unary(p)
    while p.kind() in ("Star", "Slash"):
        op = "*" if p.take()[0] == "Star" else "/"
        left = ("bin", op, left, _unary(p))
    return left


def _unary(p):
    if p.kind() == "Minus":
        p.take()
        return ("un", "neg", _primary(p))
    return _primary(p)


def _primary(p):
    k = p.kind()
    if k == "Number":
        return ("num", p.take()[1])
    if k in ("True", "False"):
        return ("bool", p.take()[0] == "True")
    if k == "LParen":
        p.take()
        e = _expr(p)
        p.take("RParen")
        return e
    if k == "Ident":
        name = p.take()[1]
        if p.kind() == "LParen":
            p.take()
            args = []
            if p.kind() != "RParen":
                args.append(_expr(p))
                while p.kind() == "Comma":
                    p.take()
                    args.append(_expr(p))
            p.take("RParen")
            return ("call", name, args)
        return ("id", name)
    if k in ("Notional", "Maturity"):
        return ("id", p.take()[1])
    tok = p.t[p.i] if p.i < len(p.t) else (None, None, 0, 0)
    raise DslError(f"parse error at {tok[2]}-{tok[3]}: expected expression, got {k}")


def _ident(p):
    k = p.kind()
    if k in ("Ident", "Notional", "Maturity"):
        return p.take()[1]
    p.take("Ident", "identifier")


def _stmt(p):
    k = p.kind()
    if k in ("Let", "Set"):
        p.take()
        name = _ident(p)
        p.take("Eq")
        return ("let" if k == "Let" else "set", name, _expr(p))
    if k == "If":
        return _if(p)
    if k in ("Pay", "Redeem"):
        p.take()
        return (k.lower(), _expr(p))
    if k == "Skip":
        p.take()
        return ("skip",)
    tok = p.t[p.i] if p.i < len(p.t) else (None, None, 0, 0)
    raise DslError(f"parse error at {tok[2]}-{tok[3]}: expected statement (let, if, pay, redeem, "
                   f"set, skip), got {k}")


def _body(p):
    p.take("Indent")
    out = []
    while not p.block_end():
        out.append(_stmt(p))
    p.take("Dedent")
    return out


def _if(p):
    p.take("If")
    cond = _expr(p)
    p.take("Then")
    then = _body(p)
    other = []
    if p.kind() == "Else":
        p.take()
        other = [_if(p)] if p.kind() == "If" else _body(p)
    return ("if", cond, then, other)


FREQ = {"Monthly": 1.0 / 12.0, "Quarterly": 0.25, "SemiAnnual": 0.5, "Annual": 1.0}


def parse(toks):
    p = _P(toks)
    p.take("Product")
    name = p.take("StringLit", "string")[1]
    p.take("Indent")
    items = []
    while not p.block_end():
        k = p.kind()
        if k in ("Notional", "Maturity"):
            p.take()
            p.take("Colon")
            items.append((k.lower(), p.take("Number", "number")[1]))
        elif k == "Underlyings":
            p.take()
            p.take("Indent")
            decls = []
            while not p.block_end():
                nm = _ident(p)
                idx = len(decls)
                if p.kind() == "Eq":
                    p.take()
                    if p.kind() not in ("Asset", "Equity", "Fx", "Commodity", "Rate"):
                        raise DslError("parse error: expected asset type")
                    p.take()
                    p.take("LParen")
                    v = p.take("Number", "number")[1]
                    idx = 0 if not v > 0 else int(v)
                    p.take("RParen")
                decls.append((nm, idx))
            p.take("Dedent")
            items.append(("underlyings", decls))
        elif k == "State":
            p.take()
            p.take("Indent")
            decls = []
            while not p.block_end():
                nm = _ident(p)
                p.take("Colon")
                tk = p.take()
                tname = {"Bool": "bool", "Float": "float"}.get(tk[0], tk[1])
                p.take("Eq")
                decls.append((nm, tname, _expr(p)))
            p.take("Dedent")
            items.append(("state", decls))
        elif k == "Schedule":
            p.take()
            fk = p.take()
            period = FREQ[fk[0]] if fk[0] in FREQ else fk[1]
            p.take("From")
            start = p.take("Number", "number")[1]
            p.take("To")
            end = p.take("Number", "number")[1]
            items.append(("schedule", period, start, end, _body(p)))
        else:
            raise DslError(f"parse error: unexpected token {k} in product body")
    p.take("Dedent")
    return name, items


# -------------------------------------------------------------- compiler --

def generate_dates(period, start, end):
    """ScheduleFreq::generate_dates (ast.rs:64-80)"""
    if period <= 0.0:
        return [end]
    dates, t = [], start
    while t <= end + 1e-12:
        dates.append(t)
        t += period
    if not dates or abs(dates[-1] - end) > 1e-12:
        dates.append(end)
    return dates


class Product:
    pass


def compile_source(src):
    """parse_and_compile (mod.rs:24-28) -> Product with resolved slots."""
    name, items = parse(tokenize(src))
    prod = Product()
    prod.name, prod.notional, prod.maturity = name, None, None
    prod.state, prod.underlyings, prod.schedules = [], [], []
    state_slots = {}
    for it in items:
        if it[0] in ("notional", "maturity"):
            if getattr(prod, it[0]) is not None:
                raise DslError(f"compile error: duplicate {it[0]} declaration")
            setattr(prod, it[0], it[1])
        elif it[0] == "underlyings":
            for nm, idx in it[1]:
                if nm in [u[0] for u in prod.underlyings]:
                    raise DslError(f"compile error: duplicate underlying '{nm}'")
                prod.underlyings.append((nm, idx))
        elif it[0] == "state":
            for nm, tname, init in it[1]:
                if nm in state_slots:
                    raise DslError(f"compile error: duplicate state variable '{nm}'")
                if init[0] == "bool" and tname in ("bool", "float"):
                    v = 1.0 if init[1] else 0.0
                elif init[0] == "num" and tname == "float":
                    v = init[1]
                elif init[0] == "num" and tname == "bool":
                    v = 1.0 if init[1] != 0.0 else 0.0
                else:
                    raise DslError(f"compile error: cannot use {init} as initial value for type "
                                   f"'{tname}'")
                state_slots[nm] = len(prod.state)
                prod.state.append(v)
    if prod.notional is None:
        raise DslError("compile error: missing notional declaration")
    if prod.maturity is None:
        raise DslError("compile error: missing maturity declaration")
    for it in items:
        if it[0] != "schedule":
            continue
        locals_ = {}
        counter = [0]

        def cexpr(e):
            k = e[0]
            if k in ("num", "bool"):
                return ("lit", float(e[1]))
            if k == "id":
                nm = e[1]
                if nm in locals_:
                    return ("local", locals_[nm])
                if nm in state_slots:
                    return ("state", state_slots[nm])
                if nm in ("notional", "observation_date", "is_final"):
                    return (nm,)
                raise DslError(f"compile error: undefined variable '{nm}'")
            if k == "bin":
                return ("bin", e[1], cexpr(e[2]), cexpr(e[3]))
            if k == "un":
                return ("un", e[1], cexpr(e[2]))
            if e[1] not in ("worst_of", "best_of", "performances", "price", "min", "max", "abs",
                            "exp", "log"):
                raise DslError(f"compile error: unknown function '{e[1]}'")
            return ("call", e[1], [cexpr(a) for a in e[2]])

        def cstmt(s):
            k = s[0]
            if k == "let":
                ex = cexpr(s[2])
                slot = counter[0]
                counter[0] += 1
                locals_[s[1]] = slot
                return ("let", slot, ex)
            if k == "set":
                if s[1] not in state_slots:
                    raise DslError(f"compile error: undefined state variable '{s[1]}'")
                return ("set", state_slots[s[1]], cexpr(s[2]))
            if k == "if":
                c = cexpr(s[1])
                return ("if", c, [cstmt(x) for x in s[2]], [cstmt(x) for x in s[3]])
            if k in ("pay", "redeem"):
                return (k, cexpr(s[1]))
            return ("skip",)

        body = [cstmt(s) for s in it[4]]
        prod.schedules.append((generate_dates(it[1], it[2], it[3]), body, counter[0]))
    return prod


def check_builtins(prod):
    """The plan builder's errors (compile_builtin, eval.rs:1643-1745)."""
    def walk(e):
        if e[0] == "call":
            fn, args = e[1], e[2]
            if fn == "performances":
                raise DslError("eval error: performances() cannot be used standalone; wrap in "
                               "worst_of() or best_of()")
            if fn in ("worst_of", "best_of"):
                if len(args) == 1 and args[0][0] == "call" and args[0][1] == "performances":
                    return
            elif fn == "price" and len(args) != 1:
                raise DslError("eval error: price() requires an asset index")
            elif fn in ("min", "max") and len(args) != 2:
                raise DslError(f"eval error: {fn}() requires 2 arguments")
            elif fn in ("abs", "exp", "log") and len(args) != 1:
                raise DslError(f"eval error: {fn}() requires 1 argument")
            for a in args:
                walk(a)
        elif e[0] == "bin":
            walk(e[2])
            walk(e[3])
        elif e[0] == "un":
            walk(e[2])

    def walk_s(s):
        if s[0] in ("let", "set"):
            walk(s[2])
        elif s[0] in ("pay", "redeem"):
            walk(s[1])
        elif s[0] == "if":
            walk(s[1])
            for x in s[2] + s[3]:
                walk_s(x)

    for _, body, _ in prod.schedules:
        for s in body:
            walk_s(s)


# ------------------------------------------------------------- evaluator --

class _End(Exception):
    pass


class _Ctx:
    __slots__ = ("spots", "s0", "notional", "date", "final", "df", "locals", "state", "pv")


def _truthy(v):
    return v != 0.0


def _ev(e, c):
    k = e[0]
    if k == "lit":
        return e[1]
    if k == "local":
        return c.locals[e[1]]
    if k == "state":
        return c.state[e[1]]
    if k == "notional":
        return c.notional
    if k == "observation_date":
        return c.date
    if k == "is_final":
        return 1.0 if c.final else 0.0
    if k == "un":
        v = _ev(e[2], c)
        return -v if e[1] == "neg" else (0.0 if _truthy(v) else 1.0)
    if k == "bin":
        a, b, op = _ev(e[2], c), _ev(e[3], c), e[1]
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        if op == "/":
            return math.nan if b == 0.0 else a / b
        if op == "==":
            return 1.0 if abs(a - b) < EPS else 0.0
        if op == "!=":
            return 1.0 if abs(a - b) >= EPS else 0.0
        if op == "<":
            return 1.0 if a < b else 0.0
        if op == "<=":
            return 1.0 if a <= b else 0.0
        if op == ">":
            return 1.0 if a > b else 0.0
        if op == ">=":
            return 1.0 if a >= b else 0.0
        if op == "and":
            return 1.0 if _truthy(a) and _truthy(b) else 0.0
        return 1.0 if _truthy(a) or _truthy(b) else 0.0
    fn, args = e[1], e[2]
    if fn in ("worst_of", "best_of"):
        worst = fn == "worst_of"
        if len(args) == 1 and args[0][0] == "call" and args[0][1] == "performances":
            v = math.inf if worst else -math.inf
            for s, s0 in zip(c.spots, c.s0):
                if s0 > 0.0:
                    perf = s / s0
                    if (perf < v) if worst else (perf > v):
                        v = perf
            return v
        vals = [_ev(a, c) for a in args]
        v = math.inf if worst else -math.inf
        for x in reversed(vals):   # popped last first
            if (x < v) if worst else (x > v):
                v = x
        return v
    if fn == "price":
        x = _ev(args[0], c)
        idx = 0 if not x >= 1.0 else (2 ** 64 - 1 if x >= 1.8446744073709552e19 else int(x))
        if idx >= len(c.spots):
            raise PricingError("NumericalError", f"eval error: asset index {idx} out of range "
                                                 f"(have {len(c.spots)} assets)")
        return c.spots[idx]
    if fn in ("min", "max"):
        a, b = _ev(args[0], c), _ev(args[1], c)
        if math.isnan(a):
            return b
        if math.isnan(b):
            return a
        return min(a, b) if fn == "min" else max(a, b)
    v = _ev(args[0], c)
    if fn == "abs":
        return abs(v)
    if fn == "exp":
        try:
            return math.exp(v)
        except OverflowError:
            return math.inf
    if v == 0.0:
        return -math.inf
    if v < 0.0 or math.isnan(v):
        return math.nan
    return math.log(v)


def _run(stmts, c):
    for s in stmts:
        k = s[0]
        if k == "let":
            c.locals[s[1]] = _ev(s[2], c)
        elif k == "set":
            c.state[s[1]] = _ev(s[2], c)
        elif k == "pay":
            c.pv += _ev(s[1], c) * c.df
        elif k == "redeem":
            c.pv += _ev(s[1], c) * c.df
            raise _End()
        elif k == "skip":
            raise _End()
        else:
            _run(s[2] if _truthy(_ev(s[1], c)) else s[3], c)


def plan(prod, n_steps):
    """Observation -> snapshot map (build_execution_plan, eval.rs:268-330)."""
    scale = n_steps / prod.maturity if prod.maturity > 0 else 0.0
    step_snap, obs = {}, []
    for dates, body, n_locals in prod.schedules:
        rows = []
        for j, t in enumerate(dates):
            step = 0
            if prod.maturity > 0:
                x = t * scale   # f64::round: half away from zero, then `as usize`
                r = math.floor(abs(x))
                if abs(x) - r >= 0.5:
                    r += 1
                r = r if x >= 0 else -r
                step = int(r) if r > 0 else 0
            step = min(step, n_steps)
            if step not in step_snap:
                step_snap[step] = len(step_snap)
            rows.append((step_snap[step], t, j + 1 == len(dates)))
        obs.append(rows)
    return step_snap, obs


def evaluate(prod, n_steps, rate, snaps, s0, max_locals):
    """evaluate_product_with_plan_in_place for one path: snaps[k] = spots of
    snapshot k."""
    _, obs = plan(prod, n_steps)
    c = _Ctx()
    c.s0, c.notional, c.pv = s0, prod.notional, 0.0
    c.state = list(prod.state)
    c.locals = [0.0] * max(1, max_locals)
    try:
        for (dates, body, _), rows in zip(prod.schedules, obs):
            for k, t, final in rows:
                c.locals = [0.0] * max(1, max_locals)
                c.spots, c.date, c.final, c.df = snaps[k], t, final, math.exp(-rate * t)
                _run(body, c)
    except _End:
        pass
    return c.pv


# ---------------------------------------------------------------- engine --

def validate_market(assets, corr):
    """MultiAssetMarket::validate (market.rs:170-215). assets: dicts with
    kind in {equity, fx, commodity, rate}."""
    if not assets:
        raise PricingError("InvalidInput", "multi-asset market requires at least one asset")
    for a in assets:
        if a["kind"] == "rate":
            if not (a["vol"] > 0.0 and math.isfinite(a["vol"])):
                raise PricingError("InvalidInput", "rate asset vol must be finite and > 0")
        else:
            if a["spot"] <= 0.0:
                raise PricingError("InvalidInput", "asset spot must be > 0")
            if not (a["vol"] > 0.0 and math.isfinite(a["vol"])):
                raise PricingError("InvalidInput", "asset vol must be finite and > 0")


def cholesky_psd(m, tol=1e-12):
    """cholesky_lower_psd (correlation.rs:362-389)"""
    n = len(m)
    l = [[0.0] * n for _ in range(n)]
    for i in range(n):
        for j in range(i + 1):
            s = m[i][j]
            for k in range(j):
                s -= l[i][k] * l[j][k]
            if i == j:
                if s < -tol:
                    return None
                l[i][j] = math.sqrt(max(s, tol))
            elif l[j][j] > tol:
                l[i][j] = s / l[j][j]
    return l


def simulate(assets, corr, rate, maturity, n_paths, n_steps, normals, step_snap):
    """Snapshots [path][snap][asset] of the reference's correlated paths
    (price_path_chunk_scalar, engine.rs:413-495). normals [path][step][asset]."""
    n = len(assets)
    dt = maturity / n_steps
    sq = math.sqrt(dt)
    chol = cholesky_psd(corr)
    if chol is None:
        raise PricingError("InvalidInput", "correlation matrix could not be factorized")
    L = np.array(chol)
    z = np.asarray(normals, dtype=np.float64).reshape(n_paths, n_steps, n)
    # correlate_normals: running sum from 0, j = 0..i
    corr_z = np.zeros_like(z)
    for i in range(n):
        acc = np.zeros(z.shape[:2])
        for j in range(i + 1):
            acc = acc + L[i, j] * z[:, :, j]
        corr_z[:, :, i] = acc
    s0 = np.array([a["spot"] for a in assets], dtype=np.float64)
    snaps = np.zeros((n_paths, len(step_snap), n))
    cur = np.tile(s0, (n_paths, 1))
    if 0 in step_snap:
        snaps[:, step_snap[0], :] = cur
    for step in range(n_steps):
        nxt = np.empty_like(cur)
        for i, a in enumerate(assets):
            zz = corr_z[:, step, i]
            if a["kind"] in ("equity", "fx"):
                mu = rate - a.get("dividend_yield", 0.0) if a["kind"] == "equity" else \
                    a["domestic_rate"] - a["foreign_rate"]
                drift = (mu - 0.5 * a["vol"] * a["vol"]) * dt
                nxt[:, i] = cur[:, i] * np.exp(drift + (a["vol"] * sq) * zz)
            elif a["kind"] == "commodity":
                ek = math.exp(-a["kappa"] * dt)
                nxt[:, i] = np.exp(a["mu"] + ek * (np.log(cur[:, i]) - a["mu"]) + (a["vol"] * sq) * zz)
            else:
                ea = math.exp(-a["mean_reversion"] * dt)
                nxt[:, i] = a["long_run_mean"] + ea * (cur[:, i] - a["long_run_mean"]) + \
                    (a["vol"] * sq) * zz
        cur = nxt
        if step + 1 in step_snap:
            snaps[:, step_snap[step + 1], :] = cur
    return snaps, s0


def price(source_or_product, assets, corr, rate, n_paths, n_steps, normals=None, seed=42):
    """price_multi_asset: (price, stderr). normals [path][step][asset] or
    None for the engine's single stream seeded by `seed`."""
    prod = compile_source(source_or_product) if isinstance(source_or_product, str) \
        else source_or_product
    validate_market(assets, corr)
    if n_paths <= 0:
        raise PricingError("InvalidInput", "num_paths must be > 0")
    if n_steps <= 0:
        raise PricingError("InvalidInput", "num_steps must be > 0")
    if prod.maturity <= 0.0:
        raise PricingError("InvalidInput", "product maturity must be > 0")
    try:
        check_builtins(prod)
    except DslError as e:
        raise PricingError("NumericalError", str(e))
    step_snap, _ = plan(prod, n_steps)
    if not step_snap:
        return 0.0, 0.0
    n = len(assets)
    if normals is None:
        from oracle import oracle as O
        normals = O.lsm_stream_normals(seed, n_paths * n_steps * n)
    snaps, s0 = simulate(assets, corr, rate, prod.maturity, n_paths, n_steps, normals, step_snap)
    max_locals = max([s[2] for s in prod.schedules] + [0])
    sum_pv = sum_sq = 0.0
    for p in range(n_paths):
        pv = evaluate(prod, n_steps, rate, [list(r) for r in snaps[p]], list(s0), max_locals)
        sum_pv += pv
        sum_sq += pv * pv
    nn = float(n_paths)
    mean = sum_pv / nn
    var = (sum_sq - sum_pv * sum_pv / nn) / (nn - 1.0) if n_paths > 1 else 0.0
    return mean, math.sqrt(var / nn)

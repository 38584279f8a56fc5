"""Writes tests/golden/second_order_spec_vectors.json: numeric vectors of the
reference's machine-readable second-order equations.

Runs only in the build container, where the reference checkout is readable.
It reads two data files of the reference (nothing of it is imported or run):
  * equation_spec_v2.yaml: the background (mu0, tau0, pi0, psi20), the
    coordinate forms of Delta_n, eth_n and eth'_n, the ordered-pair source
    D_12 and T_12, S_over_R3 and the coordinate forcing, the seven rescaled
    reconstruction equations, and the linear reduction coefficients;
  * SOURCE_TERM_LEDGER.csv: every term of D_12 (sd01-sd15), T_12 (st01-st07)
    and S / R^3 (S01-S04) with its coefficient and explicit R power.
The expressions are parsed by a small parser of the spec's notation
(implicit multiplication, [ ] as parentheses, X^sharp, bar( ), operator
application such as `Delta_1 F2`) into SymPy, and evaluated at seeded
random inputs with 50-digit arithmetic.  Jet1 tangents are the exact
directional derivatives sum_x dE/dx * xdot (SymPy differentiation; the
spec's bilinear JVP).  The generator also checks that the ledger rows sum to
the spec's D_12, T_12 and S_over_R3 exactly (two reference data files that
must agree) and that each row's R power column matches its expression.

Only numbers and symbol names are written; no spec text is copied.

Usage: python tests/golden/gen_spec_vectors.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import json
import os
import re
import sys

import numpy as np
import sympy as sp
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "second_order_spec_vectors.json")
DIGITS = 50

# ---------------------------------------------------------------- lexer
# Longest names first: regex alternation takes the first alternative that
# matches, so every name that has a shorter prefix name comes before it.
NAME = re.compile(
    r"spin_weighted_laplacian|S_over_R3|ethprime_\d|eth_\d|Delta_\d|Raise_s|Lower_s|d_T|d_R|"
    r"sqrt|sin|cos|bar|theta|mu0|tau0|pi0|psi20|Hrad|Gm|psi|gamma|sharp|forcing|"
    r"(?:Lambda|Pi)(?:_?\d)?|[DT]_mt|[FGHBCU](?:_?\d)?|[A-Za-z]")
NUMBER = re.compile(r"\d+(?:\.\d+)?")
OPERATORS = re.compile(r"^(ethprime_\d|eth_\d|Delta_\d|Raise_s|Lower_s|d_T|d_R)$")
FUNCTIONS = {"sin": sp.sin, "cos": sp.cos, "sqrt": sp.sqrt, "bar": sp.conjugate}


def tokenize(text: str):
    toks, i = [], 0
    while i < len(text):
        c = text[i]
        if c.isspace():
            i += 1
            continue
        m = NUMBER.match(text, i)
        if m:
            toks.append(("num", m.group()))
            i = m.end()
            continue
        m = NAME.match(text, i)
        if m:
            toks.append(("name", m.group()))
            i = m.end()
            continue
        if c in "+-*/^()[]":
            toks.append(("op", c))
            i += 1
            continue
        raise ValueError(f"cannot tokenize {text[i:i + 20]!r}")
    return toks


class Parser:
    """expr := term (+|- term)*;  term := unary ((*|/) unary | power)*
    (juxtaposition multiplies);  unary := (+|-) unary | power;
    power := atom (^ atom | ^sharp)?;  atom := number | name | (expr) |
    [expr] | f(expr) | OP operand."""

    def __init__(self, text: str, rename=None):
        self.t = tokenize(text)
        self.k = 0
        self.rename = rename or {}

    def peek(self):
        return self.t[self.k] if self.k < len(self.t) else (None, None)

    def take(self):
        tok = self.peek()
        self.k += 1
        return tok

    def expect(self, value):
        tok = self.take()
        if tok[1] != value:
            raise ValueError(f"expected {value}, got {tok}")

    def parse(self):
        e = self.expr()
        if self.k != len(self.t):
            raise ValueError(f"trailing tokens {self.t[self.k:]}")
        return e

    def expr(self):
        e = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            r = self.term()
            e = e + r if op == "+" else e - r
        return e

    def starts_factor(self):
        kind, v = self.peek()
        return kind in ("num", "name") or v in ("(", "[")

    def term(self):
        e = self.unary()
        while True:
            v = self.peek()[1]
            if v == "*":
                self.take()
                e = e * self.unary()
            elif v == "/":
                self.take()
                e = e / self.unary()
            elif self.starts_factor():
                e = e * self.power()
            else:
                return e

    def unary(self):
        v = self.peek()[1]
        if v == "-":
            self.take()
            return -self.unary()
        if v == "+":
            self.take()
            return self.unary()
        return self.power()

    def field_name(self, name):
        name = self.rename.get(name, name)
        m = re.fullmatch(r"([A-Za-z]+)_(\d)", name)
        if m and not OPERATORS.match(name):
            name = m.group(1) + m.group(2)
        return name

    def sharp_suffix(self):
        if self.peek()[1] == "^" and self.k + 1 < len(self.t) and self.t[self.k + 1][1] == "sharp":
            self.k += 2
            return "_s"
        return ""

    def power(self):
        base_tok = self.peek()
        base = self.atom()
        if self.peek()[1] == "^":
            if self.t[self.k + 1][1] == "sharp":
                # X^sharp on a plain field symbol
                self.k += 2
                if base_tok[0] != "name":
                    raise ValueError("^sharp on a non-field")
                return sp.Symbol(self.field_name(base_tok[1]) + "_s")
            self.take()
            return base ** self.atom()
        return base

    def atom(self):
        kind, v = self.take()
        if kind == "num":
            return sp.Rational(v)
        if v == "(":
            e = self.expr()
            self.expect(")")
            return e
        if v == "[":
            e = self.expr()
            self.expect("]")
            return e
        if kind != "name":
            raise ValueError(f"unexpected {v}")
        if v in FUNCTIONS:
            self.expect("(")
            e = self.expr()
            self.expect(")")
            return FUNCTIONS[v](e)
        if v == "spin_weighted_laplacian":
            self.expect("(")
            arg = self.take()[1]
            self.expect(")")
            return sp.Symbol(f"laplacian[{arg}]")
        if OPERATORS.match(v):
            kind2, operand = self.take()
            if kind2 != "name":
                raise ValueError(f"operator {v} without operand")
            operand = self.field_name(operand) + self.sharp_suffix()
            return sp.Symbol(f"{v}[{operand}]")
        if v == "i":
            return sp.I
        return sp.Symbol(self.field_name(v))


def parse(text: str, rename=None):
    return Parser(" ".join(str(text).split()), rename).parse()


# ---------------------------------------------------------------- numbers
CT, ST = sp.Symbol("CT"), sp.Symbol("ST")
THETA = sp.Symbol("theta")


def angles(expr):
    return expr.subs({sp.cos(THETA): CT, sp.sin(THETA): ST})


def fnum(x: float):
    return sp.Float(float(x), DIGITS)


def cnum(z: complex):
    return fnum(z.real) + sp.I * fnum(z.imag)


def evaluate(expr, values: dict) -> complex:
    subs = {sp.Symbol(k) if isinstance(k, str) else k: v for k, v in values.items()}
    got = sp.N(expr.subs(subs), DIGITS)
    free = got.free_symbols
    if free:
        raise ValueError(f"unbound symbols {free}")
    return complex(got)


def pair(z: complex):
    return [float(z.real), float(z.imag)]


def rand_c(rng):
    return complex(rng.normal(), rng.normal())


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    spec_path = os.path.join(args.reference, "equation_spec_v2.yaml")
    ledger_path = os.path.join(args.reference, "SOURCE_TERM_LEDGER.csv")
    with open(spec_path) as f:
        spec = yaml.safe_load(f)
    with open(ledger_path, newline="") as f:
        ledger = list(csv.DictReader(f))
    rng = np.random.default_rng(20260811)
    out = {"provenance": {
        "files": {os.path.basename(p): hashlib.sha256(open(p, "rb").read()).hexdigest()
                  for p in (spec_path, ledger_path)},
        "digits": DIGITS,
        "generator": "tests/golden/gen_spec_vectors.py"}}

    # ---- background (equation_spec_v2.yaml: background)
    bg_exprs = {k: angles(parse(spec["background"][k])) for k in ("mu0", "tau0", "pi0", "psi20")}

    def background(M, a, L, R, ct, st):
        vals = {"M": fnum(M), "a": fnum(a), "L": fnum(L), "R": fnum(R), CT: fnum(ct),
                ST: fnum(st)}
        return {k: evaluate(e, vals) for k, e in bg_exprs.items()}

    points = []
    for _ in range(10):
        M = float(rng.choice([1.0, 1.0, 0.8, 1.3]))
        a = float(rng.uniform(0.0, 0.99)) * M
        L = float(rng.uniform(0.8, 2.2))
        R = float(rng.uniform(0.0, 1.0)) * L * L / (M + np.sqrt(M * M - a * a))
        ct = float(rng.uniform(-1.0, 1.0))
        st = float(np.sqrt(1.0 - ct * ct))
        points.append((M, a, L, R, ct, st))
    out["background"] = [{"M": M, "a": a, "L": L, "R": R, "ct": ct, "st": st,
                          **{k: pair(v) for k, v in background(M, a, L, R, ct, st).items()}}
                         for (M, a, L, R, ct, st) in points]
    # the spec's identity Delta mu = -mu^2 in the rescaled background:
    # Delta_1 mu0 = -R mu0^2 with d_T mu0 = 0 and d_R from SymPy
    d_expr = parse(spec["rescaled_operators"]["Delta_n"]["coordinate"])
    R_, M_, L_ = sp.symbols("R M L")
    mu0 = bg_exprs["mu0"]
    delta1_mu0 = d_expr.subs({sp.Symbol("d_T[X]"): 0, sp.Symbol("d_R[X]"): sp.diff(mu0, R_),
                              sp.Symbol("X"): mu0, sp.Symbol("n"): 1})
    assert sp.simplify(delta1_mu0 + R_ * mu0 ** 2) == 0, "Delta mu = -mu^2 fails"

    # ---- Delta_n coordinate form
    vec = []
    for (M, a, L, R, ct, st) in points:
        for n in (1, 2, 3):
            X, XT, XR = rand_c(rng), rand_c(rng), rand_c(rng)
            val = evaluate(d_expr, {"M": fnum(M), "L": fnum(L), "R": fnum(R), "n": n,
                                    "X": cnum(X), "d_T[X]": cnum(XT), "d_R[X]": cnum(XR)})
            vec.append({"M": M, "L": L, "R": R, "n": n, "X": pair(X), "X_T": pair(XT),
                        "X_R": pair(XR), "Delta": pair(val)})
    out["delta_n"] = vec

    # ---- eth_n / eth'_n coordinate forms (p = s + b, q = -s + b)
    vec = []
    for prime, key, wname in ((0, "eth_n", "p"), (1, "ethprime_n", "q")):
        text, weight = spec["rescaled_operators"][key]["coordinate"].rsplit(",", 1)
        w_lhs, w_rhs = weight.split("=")
        assert w_lhs.strip() == wname
        expr = angles(parse(text).subs(sp.Symbol(wname), parse(w_rhs)))
        ang_sym = "Raise_s[X]" if prime == 0 else "Lower_s[X]"
        for (M, a, L, R, ct, st) in points[:6]:
            for s_, b_ in ((-2, -2), (-1, -1), (-2, 0), (-1, 0), (-1, 1), (0, 0), (2, 0), (1, 1),
                           (-1, -2)):
                X, XT, ANG = rand_c(rng), rand_c(rng), rand_c(rng)
                val = evaluate(expr, {"a": fnum(a), "L": fnum(L), "R": fnum(R), CT: fnum(ct),
                                      ST: fnum(st), "s": s_, "b": b_, "X": cnum(X),
                                      "d_T[X]": cnum(XT), ang_sym: cnum(ANG)})
                vec.append({"prime": prime, "spin": s_, "boost": b_, "a": a, "L": L, "R": R,
                            "ct": ct, "st": st, "X": pair(X), "X_T": pair(XT),
                            "ang": pair(ANG), "out": pair(val)})
    out["eth"] = vec

    # ---- ordered-pair source D_12, T_12 with Jet1 tangents
    src = spec["compact_ordered_pair_source"]
    D12, T12 = parse(src["D_12"]), parse(src["T_12"])
    bg_syms = {sp.Symbol(k) for k in ("mu0", "tau0", "pi0", "R")}
    inputs = sorted((D12.free_symbols | T12.free_symbols) - bg_syms, key=str)
    out["pair_source_inputs"] = [str(x) for x in inputs]

    def jet(expr, values, tangents):
        v = evaluate(expr, values)
        d = sum((evaluate(sp.diff(expr, x), values) * tangents[str(x)] for x in inputs
                 if sp.diff(expr, x) != 0), 0j)
        return v, d

    vec = []
    for (M, a, L, R, ct, st) in points:
        bgv = background(M, a, L, R, ct, st)
        vals = {str(x): rand_c(rng) for x in inputs}
        tans = {str(x): rand_c(rng) for x in inputs}
        values = {k: cnum(v) for k, v in vals.items()}
        values.update({"mu0": cnum(bgv["mu0"]), "tau0": cnum(bgv["tau0"]),
                       "pi0": cnum(bgv["pi0"]), "R": fnum(R)})
        Dv, Dt = jet(D12, values, tans)
        Tv, Tt = jet(T12, values, tans)
        vec.append({"R": R, "mu0": pair(bgv["mu0"]), "tau0": pair(bgv["tau0"]),
                    "pi0": pair(bgv["pi0"]),
                    "values": {k: pair(v) for k, v in vals.items()},
                    "tangents": {k: pair(v) for k, v in tans.items()},
                    "D": pair(Dv), "D_t": pair(Dt), "T": pair(Tv), "T_t": pair(Tt)})
    out["pair_source"] = vec

    # ---- outer source S / R^3 and the coordinate forcing
    S_expr = parse(src["S_over_R3"])
    forcing_text = src["coordinate_forcing"]
    F_expr = angles(parse(forcing_text))
    vec = []
    for (M, a, L, R, ct, st) in points:
        bgv = background(M, a, L, R, ct, st)
        vals = {"Delta_3[D_mt]": rand_c(rng), "D_mt": rand_c(rng),
                "ethprime_3[T_mt]": rand_c(rng), "T_mt": rand_c(rng)}
        values = {k: cnum(v) for k, v in vals.items()}
        values.update({"mu0": cnum(bgv["mu0"]), "tau0": cnum(bgv["tau0"]),
                       "pi0": cnum(bgv["pi0"]), "R": fnum(R)})
        S = evaluate(S_expr, values)
        Fv = evaluate(F_expr, {"L": fnum(L), "a": fnum(a), "R": fnum(R), CT: fnum(ct),
                               "S_over_R3": cnum(S)})
        vec.append({"M": M, "a": a, "L": L, "R": R, "ct": ct, "st": st, "mu0": pair(bgv["mu0"]),
                    "tau0": pair(bgv["tau0"]), "pi0": pair(bgv["pi0"]),
                    "values": {k: pair(v) for k, v in vals.items()},
                    "S_over_R3": pair(S), "forcing": pair(Fv)})
    out["outer_source"] = vec

    # ---- ledger rows: each term alone, at two radii with the background fixed
    ledger_exprs = {}
    for row in ledger:
        rid = row["id"]
        rename = {"D": "D_mt", "T": "T_mt"} if rid.startswith("S") else {}
        ledger_exprs[rid] = parse(row["correct_rescaled_ordered_pair_term"], rename)
    for prefix, full in (("sd", D12), ("st", T12), ("S", S_expr)):
        total = sum(e for k, e in ledger_exprs.items() if k.startswith(prefix))
        assert sp.simplify(sp.expand(total - full)) == 0, f"ledger {prefix} rows != spec"
    vec = []
    for row in ledger:
        rid = row["id"]
        e = ledger_exprs[rid]
        own = sorted(e.free_symbols - bg_syms, key=str)
        # R power of the row (bg held fixed): the ledger's column must match
        Rs = sp.Symbol("R")
        poly = sp.Poly(sp.expand(e), Rs)
        powers = sorted({m[0] for m in poly.monoms()})
        column = int(row["explicit_R_power_after_dividing_by_R3"])
        assert powers[0] == column, f"{rid}: R powers {powers} vs column {column}"
        coef = sp.Rational(row["coefficient"])
        (M, a, L, R, ct, st) = points[len(vec) % len(points)]
        bgv = background(M, a, L, R, ct, st)
        vals = {str(x): rand_c(rng) for x in own}
        entry = {"id": rid, "target": {"sd": "D", "st": "T", "S": "S"}[re.match(r"[a-zA-Z]+", rid).group()],
                 "coefficient": float(coef), "r_powers": powers,
                 "mu0": pair(bgv["mu0"]), "tau0": pair(bgv["tau0"]), "pi0": pair(bgv["pi0"]),
                 "values": {k: pair(v) for k, v in vals.items()}, "radii": [], "expected": []}
        for radius in (R, 2.0 * R + 0.125):
            values = {k: cnum(v) for k, v in vals.items()}
            values.update({"mu0": cnum(bgv["mu0"]), "tau0": cnum(bgv["tau0"]),
                           "pi0": cnum(bgv["pi0"]), "R": fnum(radius)})
            # all ledger rows of the same target at these inputs: only rows
            # whose every input is present contribute
            want = sum((evaluate(ledger_exprs[k], {**values, **{str(x): 0 for x in
                                                               ledger_exprs[k].free_symbols - bg_syms
                                                               if str(x) not in vals}})
                        for k in ledger_exprs if k.startswith(re.match(r"[a-zA-Z]+", rid).group())),
                       0j)
            entry["radii"].append(radius)
            entry["expected"].append(pair(want))
        vec.append(entry)
    out["ledger"] = vec

    # ---- the seven rescaled reconstruction equations: Delta_n X = R_X
    rec = spec["metric_reconstruction_rescaled"]["equations"]
    rec_exprs, rec_n = {}, {}
    for key, text in rec.items():
        m = re.fullmatch(r"Delta_(\d)_(\w+)", key)
        rec_n[m.group(2)] = int(m.group(1))
        rec_exprs[m.group(2)] = parse(text)
    rec_inputs = sorted(set().union(*(e.free_symbols for e in rec_exprs.values())) - bg_syms,
                        key=str)
    out["reconstruction_inputs"] = [str(x) for x in rec_inputs]
    out["reconstruction_n"] = rec_n
    xt = sp.Symbol("d_T[X]")
    solved = sp.solve(sp.Eq(d_expr, sp.Symbol("RHS")), xt)[0]
    vec = []
    for (M, a, L, R, ct, st) in points:
        bgv = background(M, a, L, R, ct, st)
        vals = {str(x): rand_c(rng) for x in rec_inputs}
        dr = {k: rand_c(rng) for k in rec_exprs}
        values = {k: cnum(v) for k, v in vals.items()}
        values.update({"mu0": cnum(bgv["mu0"]), "tau0": cnum(bgv["tau0"]),
                       "pi0": cnum(bgv["pi0"]), "R": fnum(R)})
        rhs = {k: evaluate(e, values) for k, e in rec_exprs.items()}
        # time derivative from Delta_n X = R_X with the coordinate form
        dt = {k: evaluate(solved, {"M": fnum(M), "L": fnum(L), "R": fnum(R), "n": rec_n[k],
                                   "X": cnum(vals[k]), "d_R[X]": cnum(dr[k]),
                                   "RHS": cnum(rhs[k])})
              for k in rec_exprs}
        vec.append({"M": M, "L": L, "R": R, "mu0": pair(bgv["mu0"]), "tau0": pair(bgv["tau0"]),
                    "pi0": pair(bgv["pi0"]), "values": {k: pair(v) for k, v in vals.items()},
                    "radial": {k: pair(v) for k, v in dr.items()},
                    "rhs": {k: pair(v) for k, v in rhs.items()},
                    "time_derivative": {k: pair(v) for k, v in dt.items()}})
    out["reconstruction"] = vec

    # ---- linear reduction coefficients (C, K, Hrad, Gm and the P_T
    # coefficients of Q and psi)
    lin = spec["linear_teukolsky_reduction"]
    coefs = {k: angles(parse(v)) for k, v in lin["coefficients"].items()}
    pt = parse(lin["equations"]["P_T"])
    coefs["q"] = sp.expand(pt).coeff(sp.Symbol("Q"))
    coefs["psi"] = sp.expand(pt).coeff(sp.Symbol("psi"))
    vec = []
    for (M, a, L, R, ct, st) in points:
        theta = float(np.arccos(ct))
        for m, s_ in ((2, -2), (-3, -2), (0, -2), (1, -1)):
            vals = {"M": fnum(M), "a": fnum(a), "L": fnum(L), "R": fnum(R), "m": m, "s": s_,
                    CT: sp.cos(fnum(theta)), ST: sp.sin(fnum(theta))}
            vec.append({"M": M, "a": a, "L": L, "R": R, "theta": theta, "m": m, "s": s_,
                        **{k: pair(evaluate(e, vals)) for k, e in coefs.items()}})
    out["linear_coefficients"] = vec

    # ---- GHP weights and falloffs (weights_and_falloffs), the ledger's
    # weights of D and T, and which reconstruction equations read each input
    out["weights"] = {f["rescaled"]: {"spin": int(f["spin"]), "boost": int(f["boost"]),
                                      "falloff": int(f["falloff"])}
                      for f in spec["weights_and_falloffs"]["fields"].values()}
    lw = {}
    for row in ledger:
        tgt = {"sd": "D", "st": "T"}.get(re.match(r"[a-zA-Z]+", row["id"]).group())
        if tgt:
            w = [int(x) for x in row["spin_boost_weight"].strip("()").split(",")]
            assert lw.setdefault(tgt, w) == w, row["id"]
    out["ledger_weights"] = lw
    out["reconstruction_readers"] = {str(x): sorted(k for k, e in rec_exprs.items()
                                                    if x in e.free_symbols)
                                     for x in rec_inputs}

    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {sum(len(v) for v in out.values() if isinstance(v, list))} vectors")


if __name__ == "__main__":
    sys.exit(main())

"""Unit strings and conversion to lumol's internal units (A, fs, u, K, rad,
particle counts): lumol-core/src/units.rs.

Unit expressions are products, quotients, integer powers and parentheses of
the units of the conversion table (units.rs:30-82), parsed with the
operator precedence of units.rs:148-241 (^ above * and /, all
left-associative) and evaluated left to right as the reference does.
"""
import math

from .consts import AVOGADRO_NUMBER, BOHR_RADIUS

# units.rs:28
U_IN_KG = 1.660538782e-27

# units.rs:30-82
CONVERSION_FACTORS = {
    # distances
    "A": 1.0, "nm": 10.0, "pm": 1e-2, "fm": 1e-5, "m": 1e10, "bohr": BOHR_RADIUS,
    # time
    "fs": 1.0, "ps": 1e3, "ns": 1e6,
    # mass
    "u": 1.0, "Da": 1.0, "kDa": 1.0, "g": 1e-3 / U_IN_KG, "kg": 1.0 / U_IN_KG,
    # temperature, quantity of matter
    "K": 1.0, "mol": AVOGADRO_NUMBER,
    # angles
    "rad": 1.0, "deg": math.pi / 180.0,
    # energy
    "J": 1e-10 / U_IN_KG, "kJ": 1e-7 / U_IN_KG, "kcal": 4.184 * 1e-7 / U_IN_KG,
    "eV": 1.60217653e-19 * 1e-10 / U_IN_KG, "H": 4.35974417e-18 * 1e-10 / U_IN_KG,
    "Ry": 4.35974417e-18 / 2.0 * 1e-10 / U_IN_KG,
    # force
    "N": 1e-20 / U_IN_KG,
    # pressure
    "Pa": 1e-40 / U_IN_KG, "kPa": 1e-37 / U_IN_KG, "MPa": 1e-34 / U_IN_KG, "bar": 1e-35 / U_IN_KG,
    "atm": 101325.0 * 1e-40 / U_IN_KG,
}


class UnitError(ValueError):
    """A unit string that does not parse (units.rs ParseError)."""


_OPS = {"*": 10, "/": 10, "^": 20}


def _tokens(text):
    out, cur = [], ""
    for c in text:
        if c in "*/^()":
            if cur:
                out.append(cur)
                cur = ""
            out.append(c)
        elif not c.isspace():
            cur += c
    if cur:
        out.append(cur)
    return out


def _postfix(tokens):
    """Shunting-yard to reverse polish notation (all operators left-associative)."""
    ops, out = [], []
    for t in tokens:
        if t in _OPS:
            while ops and ops[-1] in _OPS and _OPS[t] <= _OPS[ops[-1]]:
                out.append(ops.pop())
            ops.append(t)
        elif t == "(":
            ops.append(t)
        elif t == ")":
            while ops and ops[-1] != "(":
                out.append(ops.pop())
            if not ops:
                raise UnitError("Parentheses are not equilibrated.")
            ops.pop()
        else:
            out.append(t)
    while ops:
        t = ops.pop()
        if t in "()":
            raise UnitError("Parentheses are not equilibrated.")
        out.append(t)
    return out


def powi(x, n):
    """x^n for an integer n by binary exponentiation (the integer power that
    Rust's f64::powi lowers to), so that repeated products round as there."""
    recip = n < 0
    n = abs(n)
    r = 1.0
    while True:
        if n & 1:
            r *= x
        n //= 2
        if n == 0:
            break
        x *= x
    return 1.0 / r if recip else r


def _read(stream):
    if not stream:
        raise UnitError("Malformed expression: missing a value")
    t = stream.pop()
    if t in ("*", "/"):
        rhs = _read(stream)
        lhs = _read(stream)
        return lhs * rhs if t == "*" else lhs / rhs
    if t == "^":
        if not stream or stream[-1] in _OPS:
            raise UnitError("Malformed expression: Missing value after '^'")
        p = stream.pop()
        try:
            n = int(p)
        except ValueError:
            raise UnitError(f"invalid power '{p}'") from None
        if not (p.lstrip("+-").isdigit()):
            raise UnitError(f"invalid power '{p}'")
        return powi(_read(stream), n)
    if t not in CONVERSION_FACTORS:
        raise UnitError(f"Unit '{t}' not found.")
    return CONVERSION_FACTORS[t]


def parse(unit):
    """Value of a unit expression in internal units (UnitExpr::parse + eval,
    units.rs:258-290)."""
    stream = _postfix(_tokens(unit))
    if not stream:
        raise UnitError("Malformed expression: missing a value")
    value = _read(stream)
    if stream:
        raise UnitError("Malformed expression: remaining values after the end of the unit: " + " ".join(stream))
    return value


def from_(value, unit):
    """units::from (units.rs:358-361): `value` in `unit` -> internal units."""
    return parse(unit) * value


def from_str(text):
    """units::from_str (units.rs:370-381): "<value> <unit expression>"."""
    parts = text.split()
    if not parts:
        raise UnitError("empty value")
    unit = " ".join(parts[1:])
    factor = parse(unit) if unit else 1.0
    try:
        value = float(parts[0])
    except ValueError:
        raise UnitError(f"invalid number '{parts[0]}'") from None
    return factor * value


def to(value, unit):
    """units::to (units.rs:390-393): internal units -> `unit`."""
    return value / parse(unit)

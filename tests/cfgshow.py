"""Builders for the reference's CFG values and a printer of GHC's derived
Show text (test infrastructure).

The HIP backend's offload pass (Language.Hydrangea.CodegenHIP) writes each
offloaded loop as the `show` of a HipSite record holding the reference's own
CFG values (src/Language/Hydrangea/CFG.hs, CFGCore.hs derive Show).  GHC is
not available here, so the example programs' sites are built with these
helpers and printed with the rules of derived Show:

  * constructor application `Con a b`, arguments at precedence 11 (compound
    arguments parenthesised), the whole parenthesised above precedence 10;
  * records `Con {f = v, g = w}`, fields at precedence 0, parenthesised at
    precedence 11;
  * negative numbers parenthesised above precedence 6;
  * Double as GHC prints it: the shortest round-trip digits, fixed notation
    for 0.1 <= |x| < 10^7, else `d.ddde<n>` (Numeric.showFloat);
  * Data.Map as `fromList [(k,v),...]` in key order.
"""
import math

__all__ = ["C", "R", "show"]


class C:
    """Constructor application (nullary when no arguments)."""
    __slots__ = ("name", "args")

    def __init__(self, name, *args):
        self.name = name
        self.args = args

    def __repr__(self):
        return show(self)


class R:
    """Record construction; fields in declaration order."""
    __slots__ = ("name", "fields")

    def __init__(self, name, fields):
        self.name = name
        self.fields = list(fields)

    def get(self, f):
        for k, v in self.fields:
            if k == f:
                return v
        raise KeyError(f)

    def __repr__(self):
        return show(self)


def hs_double(x):
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    r = repr(abs(x))  # shortest round-trip digits
    mant, _, exp = r.partition("e")
    digits = mant.replace(".", "")
    point = mant.index(".") if "." in mant else len(mant)
    e10 = point + (int(exp) if exp else 0)  # value = 0.digits * 10^e10
    stripped = digits.lstrip("0")
    e10 -= len(digits) - len(stripped)
    digits = stripped.rstrip("0") or "0"
    if 0.1 <= abs(x) < 1e7:
        if e10 <= 0:
            return sign + "0." + "0" * (-e10) + digits
        ip = digits[:e10].ljust(e10, "0")
        fp = digits[e10:] or "0"
        return sign + ip + "." + fp
    return sign + digits[0] + "." + (digits[1:] or "0") + "e" + str(e10 - 1)


def hs_string(s):
    out = ['"']
    for ch in s:
        o = ord(ch)
        if ch == '"':
            out.append('\\"')
        elif ch == "\\":
            out.append("\\\\")
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\t":
            out.append("\\t")
        elif 32 <= o < 127:
            out.append(ch)
        else:
            out.append("\\" + str(o))
    out.append('"')
    return "".join(out)


def show(v, d=0):
    if isinstance(v, bool):
        return "True" if v else "False"
    if isinstance(v, int):
        s = str(v)
        return "(" + s + ")" if v < 0 and d > 6 else s
    if isinstance(v, float):
        s = hs_double(v)
        return "(" + s + ")" if s.startswith("-") and d > 6 else s
    if isinstance(v, str):
        return hs_string(v)
    if isinstance(v, list):
        return "[" + ",".join(show(x) for x in v) + "]"
    if isinstance(v, tuple):
        return "(" + ",".join(show(x) for x in v) + ")"
    if isinstance(v, dict):
        s = "fromList " + show(sorted(v.items(), key=lambda kv: kv[0].encode()))
        return "(" + s + ")" if d > 10 else s
    if isinstance(v, C):
        if not v.args:
            return v.name
        s = v.name + " " + " ".join(show(a, 11) for a in v.args)
        return "(" + s + ")" if d > 10 else s
    if isinstance(v, R):
        s = v.name + " {" + ", ".join(k + " = " + show(x) for k, x in v.fields) + "}"
        return "(" + s + ")" if d >= 11 else s
    raise TypeError("cannot show %r" % (v,))

"""Serial interpreter of the reference's CFG loops - TEST INFRASTRUCTURE ONLY.

The oracle of the kernels hydgen generates: it runs an offloaded loop the way
the C the reference's C backend prints for it runs (one thread, source order),
so that the generated kernels can be compared with it on the same inputs.
Only tests/ use it, as the checker.

Semantics restated from src/Language/Hydrangea/CodegenC.hs:
  * genRHS / genBinOp / genUnOp (:2293-2525): fp64 + - * / are single IEEE
    operations (Python floats are IEEE binary64); sqrt is the correctly
    rounded square root; expf -> libm exp (math.exp calls the host libm); abs_f
    -> fabs; floor/ceil; float_of -> (double) of an int64 (round to nearest);
    int_of -> (int64_t) of a double (truncation); comparisons, && and || give
    0/1 (bools are int64, :2016);
  * integer + - * wrap mod 2^64 (the oracle's C is built with -fwrapv), / and
    % truncate toward zero (C semantics, :2473-2474);
  * loops run from their origin over their trip count, several iterators
    nested with the first outermost (collapse order, :1005-1100); a loop with
    a ReductionSpec starts its accumulator at rsInit unless it was assigned
    before (:1305-1330);
  * shape helpers as runtime/hydrangea_runtime.h:237-290 (row-major, last
    axis fastest).
Values: arrays are numpy int64/float64 arrays (written in place), tuples are
Python tuples of ints, scalars Python int/float.
"""
import math

import numpy as np

M64 = (1 << 64) - 1


def wrap(x):
    x &= M64
    return x - (1 << 64) if x >> 63 else x


def c_div(a, b):
    q = abs(a) // abs(b)
    return wrap(-q if (a < 0) != (b < 0) else q)


def c_mod(a, b):
    return wrap(a - c_div(a, b) * b)


def f_div(a, b):
    if b == 0.0:
        return float(np.float64(a) / np.float64(b))
    return a / b


def f_sqrt(a):
    if math.isnan(a):
        return a
    if a < 0.0:
        return float("nan")
    return math.sqrt(a)  # sqrt(-0.0) is -0.0, as in C


def f_exp(a):
    if math.isnan(a):
        return a
    try:
        return math.exp(a)
    except OverflowError:
        return float("inf")


def f_floor(a):
    return a if (math.isinf(a) or math.isnan(a)) else float(math.floor(a)) if a != 0.0 else a


def f_ceil(a):
    if math.isinf(a) or math.isnan(a) or a == 0.0:
        return a
    r = float(math.ceil(a))
    return -0.0 if r == 0.0 and a < 0 else r


def int_of(a):
    if math.isnan(a) or math.isinf(a) or not (-2.0 ** 63 <= a < 2.0 ** 63):
        raise ValueError("int_of outside int64 (undefined in C)")
    return int(a)


BIN = {
    "CAdd": lambda a, b: wrap(a + b), "CSub": lambda a, b: wrap(a - b), "CMul": lambda a, b: wrap(a * b),
    "CDiv": c_div, "CMod": c_mod,
    "CEq": lambda a, b: int(a == b), "CNeq": lambda a, b: int(a != b), "CLt": lambda a, b: int(a < b),
    "CLe": lambda a, b: int(a <= b), "CGt": lambda a, b: int(a > b), "CGe": lambda a, b: int(a >= b),
    "CAnd": lambda a, b: int(bool(a) and bool(b)), "COr": lambda a, b: int(bool(a) or bool(b)),
    "CAddF": lambda a, b: a + b, "CSubF": lambda a, b: a - b, "CMulF": lambda a, b: a * b, "CDivF": f_div,
    "CEqF": lambda a, b: int(a == b), "CNeqF": lambda a, b: int(a != b), "CLtF": lambda a, b: int(a < b),
    "CLeF": lambda a, b: int(a <= b), "CGtF": lambda a, b: int(a > b), "CGeF": lambda a, b: int(a >= b),
}


def _neg(a):
    return -a if isinstance(a, float) else wrap(-a)


UN = {
    "CNot": lambda a: int(not a), "CNeg": _neg, "CSqrt": f_sqrt, "CExpF": f_exp, "CAbsF": abs,
    "CFloorF": f_floor, "CCeilF": f_ceil, "CFloatOf": float, "CIntOf": int_of,
}


def shape_size(t):
    s = 1
    for e in t:
        s = wrap(s * e)
    return s


def flat_to_nd(flat, shp):
    idx = [0] * len(shp)
    rem = flat
    for i in range(len(shp) - 1, -1, -1):
        idx[i] = c_mod(rem, shp[i])
        rem = c_div(rem, shp[i])
    return tuple(idx)


def nd_to_flat(nd, shp):
    flat, stride = 0, 1
    for i in range(len(shp) - 1, -1, -1):
        flat = wrap(flat + nd[i] * stride)
        stride = wrap(stride * shp[i])
    return flat


class Interp:
    """Runs one HipSite's loop over `env` (name -> value).  `types` (the
    site's hsTypes) gives the element type of the arrays RArrayAlloc makes:
    zero-filled, as hyd_array_alloc's calloc (runtime/hydrangea_runtime.h:
    292-310)."""

    def __init__(self, env, types=None):
        self.env = env
        self.types = types or {}

    def alloc(self, var, shape):
        t = self.types.get(var)
        elem = t.args[0].name if t is not None and t.name == "CTArray" else "CTInt64"
        return np.zeros(tuple(shape), dtype=np.float64 if elem == "CTDouble" else np.int64)

    def atom(self, a):
        k = a.name
        if k == "AVar":
            return self.env[a.args[0]]
        if k == "AInt":
            return a.args[0]
        if k == "AFloat":
            return float(a.args[0])
        if k == "ABool":
            return 1 if a.args[0] else 0
        if k == "AUnit":
            return 0
        raise ValueError("atom " + k)

    def index(self, e):
        k = e.name
        if k == "IVar":
            return self.env[e.args[0]]
        if k == "IConst":
            return e.args[0]
        a = self.index(e.args[0])
        b = self.index(e.args[1])
        if k == "IAdd":
            return wrap(a + b)
        if k == "ISub":
            return wrap(a - b)
        if k == "IMul":
            return wrap(a * b)
        if k == "IDiv":
            return c_div(a, b)
        raise ValueError("index expression " + k)

    def rhs(self, r):
        k, A = r.name, r.args
        if k == "RAtom":
            return self.atom(A[0])
        if k == "RBinOp":
            return BIN[A[0].name](self.atom(A[1]), self.atom(A[2]))
        if k == "RUnOp":
            return UN[A[0].name](self.atom(A[1]))
        if k == "RArrayLoad":
            v = self.atom(A[0]).reshape(-1)[self.atom(A[1])]
            return float(v) if v.dtype == np.float64 else int(v)
        if k == "RTuple":
            return tuple(self.atom(a) for a in A[0])
        if k == "RProj":
            return self.atom(A[1])[A[0]]
        if k == "RArrayShape":
            return tuple(self.atom(A[0]).shape)
        if k == "RShapeSize":
            return shape_size(self.atom(A[0]))
        if k == "RShapeInit":
            return tuple(self.atom(A[0])[:-1])
        if k == "RShapeLast":
            t = self.atom(A[0])
            return t[-1] if t else 0
        if k == "RFlatToNd":
            return flat_to_nd(self.atom(A[0]), self.atom(A[1]))
        if k == "RNdToFlat":
            if A[0].name == "AUnit":
                return 0
            return nd_to_flat(self.atom(A[0]), self.atom(A[1]))
        if k == "R2DToFlat":
            return wrap(self.atom(A[0]) * self.atom(A[1]))
        raise ValueError("rhs " + k)

    def stmts(self, ss, assigned):
        for s in ss:
            k = s.name
            if k == "SAssign":
                if s.args[1].name == "RArrayAlloc":
                    self.env[s.args[0]] = self.alloc(s.args[0], self.atom(s.args[1].args[0]))
                else:
                    self.env[s.args[0]] = self.rhs(s.args[1])
                assigned.add(s.args[0])
            elif k == "SArrayWrite":
                arr = self.atom(s.args[0])
                arr.reshape(-1)[self.atom(s.args[1])] = self.atom(s.args[2])
            elif k == "SIf":
                if self.stmts(s.args[1] if self.atom(s.args[0]) else s.args[2], assigned):
                    return True  # a break inside the branch leaves the enclosing loop
            elif k == "SLoop":
                self.loop(s, assigned)
            elif k == "SBreak":
                return True
            else:
                raise ValueError("stmt " + k)
        return False

    def loop(self, s, assigned):
        spec, body = s.args
        iters = spec.get("lsIters")
        bounds = [self.index(b) for b in spec.get("lsBounds")]
        origins = [self.index(o) for o in spec.get("lsOrigins")] or [0] * len(iters)
        red = spec.get("lsRed")
        if red.name == "Just":
            acc = red.args[0].get("rsAccVar")
            if acc not in assigned:
                self.env[acc] = self.index(red.args[0].get("rsInit"))
                assigned.add(acc)

        # the iterators as nested C for loops: a break leaves the innermost
        def run(d):
            last = d == len(iters) - 1
            for v in range(origins[d], origins[d] + bounds[d]):
                self.env[iters[d]] = v
                if last:
                    if self.stmts(body, assigned):
                        return
                else:
                    run(d + 1)

        if iters:
            run(0)
        else:
            self.stmts(body, assigned)


def run_site(site, env):
    """Executes site.hsLoop over env (arrays written in place).  Returns the
    site's result (a reduction's accumulator, an iterate's final array), else
    None."""
    interp = Interp(env, site.get("hsTypes"))
    assigned = set(name for name, _ in site.get("hsArgs"))
    interp.stmts([site.get("hsLoop")], assigned)
    res = site.get("hsResult")
    return env[res.args[0][0]] if res.name == "Just" else None

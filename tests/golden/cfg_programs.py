"""The offload sites of the HIP backend's example programs (test fixtures).

Each program is the list of HipSite records Language.Hydrangea.CodegenHIP
emits for it: the loops the reference's lowering produces for the program's
fused combinators, after the Metal-style pipeline (optimize + parallelize,
CFGPipeline.hs:224-226) has marked them.  GHC is not available here, so the
loops are built by hand, each following the lowering it cites, and printed as
GHC's derived Show prints them (tests/cfgshow.py).  make_cfg_sites.py writes
them to tests/golden/cfg/*.sites, which the build compiles with hydgen.

Programs (loop shapes from src/Language/Hydrangea/Lowering.hs):
  map_chain    the 8-op zipwith chain of the dot + map-chain workload, one
               fused LoopMap loop (zipwith lowering; `if` as SIf, :978-1001 of
               CodegenC.hs)
  map_chain2   a second elementwise chain over fp64 and int64 inputs with two
               outputs: floor/ceil, int_of/float_of, C `/` and `%` on negative
               values, wrapping int64 products, comparisons, && and negation
  dot          reduce (+.) 0.0 (zipwith (*.) xs ys): a LoopReductionWrapper of
               trip count 1 around the reduction loop (:1150-1182)
  reduce_i64   reduce (+) over x*x and reduce (*) over (x % 5 + 3) on int64
               (wrapping), and reduce (*.) on fp64
  jacobi_step  one clamp-stencil heat step split into the 9 region loops of
               lowerSplitStencil2D (:2562-2665, :2431-2480, :2215-2340)
  wrap_stencil a 5-point stencil with the wrap boundary: the generic flat loop
               with genBnd2D's BWrap index arithmetic (:1997-2030, :2742-2880)
  mirror_1d    a 1-D 3-point stencil with the mirror boundary (genBnd1D)
  histogram    the fused weighted histogram scatter (+) (fill [B] 0) (map (x % B)
               data) (map (x * x) data) (:1408-1430, Fusion.hs:2010-2030)
  splat_f64    scatter_guarded (+.) with generated route/value/guard
               functions (:1432-1470): order-exact fp64 accumulation
  row_sums     reduce (+.) over the rows of a 2-D array: LoopMapReduction
               around a LoopReduction (:724-760)
  gather       out[i] = src[perm[i]] (:1860-1898)
  map3d        a 3-index generate over an N-D shape (RTuple + RNdToFlat)
  heat_iterate iterate of the clamp heat step (the LoopIterate loop around
               jacobi_step's region loops, a small-grid branch, u = out)
  wide_iterate iterate of a radius-2 clamp stencil (4th-order heat on a
               9-point cross): 25 region loops, a small-grid branch
  interior_iterate  iterate of a step that writes only the interior of its
               fresh (zeroed) array
  int_iterate  iterate of an int64 clamp stencil (wrapping sums, C %)
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))

from cfgshow import C, R  # noqa: E402

NOTHING = C("Nothing")
T_I, T_F, T_B, T_T = C("CTInt64"), C("CTDouble"), C("CTBool"), C("CTTuple")


def arr(t):
    return C("CTArray", t)


def atomize(a):
    if isinstance(a, C):
        return a
    if isinstance(a, bool):
        return C("ABool", a)
    if isinstance(a, str):
        return C("AVar", a)
    if isinstance(a, int):
        return C("AInt", a)
    if isinstance(a, float):
        return C("AFloat", a)
    raise TypeError(a)


def iexpr(b):
    if isinstance(b, C):
        return b
    return C("IVar", b) if isinstance(b, str) else C("IConst", b)


def assign(v, rhs):
    return C("SAssign", v, rhs)


def atom(a):
    return C("RAtom", atomize(a))


def load(a, i):
    return C("RArrayLoad", atomize(a), atomize(i))


def bop(op, a, b):
    return C("RBinOp", C(op), atomize(a), atomize(b))


def uop(op, a):
    return C("RUnOp", C(op), atomize(a))


def write(a, i, v):
    return C("SArrayWrite", atomize(a), atomize(i), atomize(v))


def sif(c, t, e=()):
    return C("SIf", atomize(c), list(t), list(e))


def pspec(strategy="ParallelGeneric"):
    return R("ParallelSpec", [("psStrategy", C(strategy)), ("psPolicy", NOTHING), ("psSimdLen", NOTHING)])


def par(strategy="ParallelGeneric"):
    return C("Parallel", pspec(strategy))


SERIAL = C("Serial")


def loop(iters, bounds, body, exec_=SERIAL, red=None, role="LoopMap", origins=None):
    spec = R("LoopSpec", [
        ("lsIters", list(iters)),
        ("lsBounds", [iexpr(b) for b in bounds]),
        ("lsExec", exec_),
        ("lsRed", NOTHING if red is None else
         C("Just", R("ReductionSpec", [("rsAccVar", red[0]), ("rsInit", iexpr(red[1])), ("rsRedop", C(red[2]))]))),
        ("lsRole", C(role)),
        ("lsOrigins", [iexpr(o) for o in origins] if origins is not None else []),
    ])
    return C("SLoop", spec, list(body))


# ---- types of the loop's variables (the table recoverProcTypeEnv gives) -----

FLOAT_BIN = {"CAddF", "CSubF", "CMulF", "CDivF"}
CMP = {"CEq", "CNeq", "CLt", "CLe", "CGt", "CGe", "CAnd", "COr", "CEqF", "CNeqF", "CLtF", "CLeF", "CGtF", "CGeF"}


def infer_types(args, lp, omit=(), extra=()):
    env = dict(args)
    env.update(dict(extra))  # arrays RArrayAlloc creates: their element type is not in the RHS

    def atype(a):
        if a.name == "AVar":
            return env[a.args[0]]
        return {"AInt": T_I, "AFloat": T_F, "ABool": T_B, "AUnit": T_I}[a.name]

    def rtype(r):
        k, A = r.name, r.args
        if k == "RAtom":
            return atype(A[0])
        if k == "RBinOp":
            op = A[0].name
            return T_B if op in CMP else T_F if op in FLOAT_BIN else T_I
        if k == "RUnOp":
            op = A[0].name
            return {"CNot": T_B, "CIntOf": T_I, "CNeg": atype(A[1])}.get(op, T_F)
        if k == "RArrayLoad":
            return atype(A[0]).args[0]
        if k in ("RTuple", "RArrayShape", "RShapeInit", "RFlatToNd"):
            return T_T
        return T_I

    def walk(ss):
        for s in ss:
            if s.name == "SAssign" and s.args[0] not in env:
                env[s.args[0]] = rtype(s.args[1])
            elif s.name == "SLoop":
                for it in s.args[0].get("lsIters"):
                    env[it] = T_I
                walk(s.args[1])
            elif s.name == "SIf":
                walk(s.args[1])
                walk(s.args[2])

    walk([lp])
    return {k: v for k, v in env.items() if k not in omit}


def site(name, args, lp, result=None, omit_types=(), extra_types=()):
    types = infer_types(args, lp, omit_types, extra_types)
    return R("HipSite", [
        ("hsName", name),
        ("hsArgs", [(v, t) for v, t in args]),
        ("hsResult", NOTHING if result is None else C("Just", result)),
        ("hsTypes", types),
        ("hsLoop", lp),
    ])


# ---- programs ---------------------------------------------------------------

def map_chain():
    body = [
        assign("x", load("xs", "i")), assign("y", load("ys", "i")),
        assign("t1", bop("CMulF", 2.0, "x")), assign("t2", bop("CAddF", "t1", "y")),
        assign("t3", uop("CAbsF", "t2")), assign("t4", uop("CSqrt", "t3")),
        assign("t5", bop("CMulF", "t4", "x")), assign("t6", bop("CSubF", "t5", "y")),
        assign("c", bop("CGtF", "t6", 0.0)),
        sif("c", [assign("t7", atom("t6"))], [assign("t7", atom(0.0))]),
        assign("t8", bop("CMulF", "t7", 1.0e-3)), assign("t9", uop("CExpF", "t8")),
        write("out", "i", "t9"),
    ]
    lp = loop(["i"], ["n"], body, par(), role="LoopMap")
    return [site("hydk_map_chain_0", [("n", T_I), ("out", arr(T_F)), ("xs", arr(T_F)), ("ys", arr(T_F))], lp)]


def map_chain2():
    body = [
        assign("x", load("xs", "i")), assign("k", load("ks", "i")),
        assign("a", bop("CMulF", "x", "scale")), assign("b", uop("CFloorF", "a")),
        assign("c", uop("CIntOf", "b")), assign("d", bop("CMod", "c", "m")),
        assign("e", bop("CDiv", "k", 7)), assign("f", bop("CSub", "d", "e")),
        assign("g", bop("CMul", "f", "k")), assign("h", uop("CFloatOf", "d")),
        assign("p", bop("CDivF", "h", "x")), assign("q", uop("CCeilF", "p")),
        assign("r", uop("CNeg", "q")), assign("lt", bop("CLtF", "x", 0.0)),
        sif("lt", [assign("s", atom("r"))], [assign("s", atom("q"))]),
        assign("eq", bop("CEq", "k", 0)), assign("ne", uop("CNot", "eq")),
        assign("both", bop("CAnd", "lt", "ne")),
        sif("both", [assign("g2", uop("CNeg", "g"))], [assign("g2", atom("g"))]),
        assign("t", bop("CSubF", "s", 1.5)), assign("ta", uop("CAbsF", "t")), assign("u", uop("CSqrt", "ta")),
        write("of", "i", "u"), write("oi", "i", "g2"),
    ]
    lp = loop(["i"], ["n"], body, par(), role="LoopMap")
    args = [("ks", arr(T_I)), ("m", T_I), ("n", T_I), ("of", arr(T_F)), ("oi", arr(T_I)), ("scale", T_F),
            ("xs", arr(T_F))]
    # a few locals left out of the type table: hydgen infers them
    return [site("hydk_map_chain2_0", args, lp, omit_types=("g2", "ta", "both"))]


def dot():
    inner = loop(["k"], ["red_dim"], [
        assign("x", load("xs", "k")), assign("y", load("ys", "k")),
        assign("elem", bop("CMulF", "x", "y")), assign("acc", bop("CAddF", "acc", "elem")),
    ], par(), red=("acc", 0, "RAdd"), role="LoopReduction")
    lp = loop(["j"], [1], [inner], SERIAL, role="LoopReductionWrapper")
    return [site("hydk_dot_0", [("acc", T_F), ("red_dim", T_I), ("xs", arr(T_F)), ("ys", arr(T_F))], lp,
                 result=("acc", T_F))]


def reduce_i64():
    s1 = loop(["k"], ["n"], [
        assign("x", load("ks", "k")), assign("sq", bop("CMul", "x", "x")), assign("acc", bop("CAdd", "acc", "sq")),
    ], par(), red=("acc", 0, "RAdd"), role="LoopReduction")
    s2 = loop(["k"], ["n"], [
        assign("x", load("ks", "k")), assign("r", bop("CMod", "x", 5)), assign("f", bop("CAdd", "r", 3)),
        assign("prod", bop("CMul", "prod", "f")),
    ], par(), red=("prod", 1, "RMul"), role="LoopReduction")
    s3 = loop(["k"], ["n"], [
        assign("x", load("xs", "k")), assign("s", bop("CMulF", "x", 1.0e-3)), assign("f", bop("CAddF", 1.0, "s")),
        assign("pf", bop("CMulF", "pf", "f")),
    ], par(), red=("pf", 1, "RMul"), role="LoopReduction")
    return [site("hydk_sum_sq_i64", [("acc", T_I), ("ks", arr(T_I)), ("n", T_I)], s1, result=("acc", T_I)),
            site("hydk_prod_i64", [("ks", arr(T_I)), ("n", T_I), ("prod", T_I)], s2, result=("prod", T_I)),
            site("hydk_prod_f64", [("n", T_I), ("pf", T_F), ("xs", arr(T_F))], s3, result=("pf", T_F))]


# The 5-point heat body: a 0 0 +. 0.1 *. (a (-1) 0 +. a 1 0 +. a 0 (-1) +. a 0 1 -. 4.0 *. a 0 0)
HEAT_ACCESSES = [((-1, 0), "a_n"), ((1, 0), "a_s"), ((0, -1), "a_w"), ((0, 1), "a_e"), ((0, 0), "a_c")]


def heat_body(res):
    return [
        assign("t1", bop("CAddF", "a_n", "a_s")), assign("t2", bop("CAddF", "t1", "a_w")),
        assign("t3", bop("CAddF", "t2", "a_e")), assign("t4", bop("CMulF", 4.0, "a_c")),
        assign("t5", bop("CSubF", "t3", "t4")), assign("t6", bop("CMulF", 0.1, "t5")),
        assign(res, bop("CAddF", "a_c", "t6")),
    ]


def int_body(res):
    """(a(-1)0 + a 1 0 + a 0(-1) + a 0 1 + 3 * a 0 0) % 1000003 on int64"""
    return [
        assign("k1", bop("CAdd", "a_n", "a_s")), assign("k2", bop("CAdd", "k1", "a_w")),
        assign("k3", bop("CAdd", "k2", "a_e")), assign("k4", bop("CMul", 3, "a_c")),
        assign("k5", bop("CAdd", "k3", "k4")), assign(res, bop("CMod", "k5", 1000003)),
    ]


def jacobi_step(body_fn=None):
    """lowerSplitStencil2D's 'large' case (n0, n1 >= 2) of the clamp heat step:
    rows and columns each split into interior (offset 1, n - 2 iterations),
    left (offset 0, one iteration) and right (offset bd_start = n - 1, one
    iteration) segments (genericStencilSegments1D); one 2-index LoopMap loop per
    (row segment, column segment) (mkStencilLoop2D); clamped indices folded per
    segment (specializeClampIndex1D); flat indices through RTuple/RNdToFlat with
    the input's shape (buildFlatIndex2D)."""
    segs0 = [("interior", C("IAdd", C("IVar", "n0"), C("IConst", -2)), 1),
             ("left", 1, 0), ("right", 1, "bd_start_r")]
    segs1 = [("interior", C("IAdd", C("IVar", "n1"), C("IConst", -2)), 1),
             ("left", 1, 0), ("right", 1, "bd_start_c")]
    sites = []
    for rk, rbound, roff in segs0:
        for ck, cbound, coff in segs1:
            body = []
            used = {"shp", "u", "out"}

            def shift(it, off, tag):
                if off == 0:
                    return it
                v = "%s_%s" % (it, tag)
                body.append(assign(v, bop("CAdd", it, off)))
                if isinstance(off, str):
                    used.add(off)
                return v

            ra = shift("i", roff, "r")
            ca = shift("j", coff, "c")

            def clamp(kind, base, off, nm1, tag):
                if kind == "interior" or (kind == "left" and off >= 0) or (kind == "right" and off < 1):
                    if off == 0:
                        return base
                    v = "%s_%s" % (base, tag)
                    body.append(assign(v, bop("CAdd", base, off)))
                    return v
                if kind == "left":
                    return 0
                used.add(nm1)
                return nm1

            for (dr, dc), res in HEAT_ACCESSES:
                tag = "%s%s" % (("m%d" % -dr) if dr < 0 else "p%d" % dr, ("m%d" % -dc) if dc < 0 else "p%d" % dc)
                r = clamp(rk, ra, dr, "n0m1", "r" + tag)
                c = clamp(ck, ca, dc, "n1m1", "c" + tag)
                body.append(assign("idx_" + res, C("RTuple", [atomize(r), atomize(c)])))
                body.append(assign("flat_" + res, C("RNdToFlat", atomize("idx_" + res), atomize("shp"))))
                body.append(assign(res, load("u", "flat_" + res)))
            body += (body_fn or heat_body)("val")
            body.append(assign("idx_o", C("RTuple", [atomize(ra), atomize(ca)])))
            body.append(assign("flat_o", C("RNdToFlat", atomize("idx_o"), atomize("shp"))))
            body.append(write("out", "flat_o", "val"))
            if rk == "interior":
                used.add("n0")
            if ck == "interior":
                used.add("n1")
            lp = loop(["i", "j"], [rbound, cbound], body, par(), role="LoopMap")
            types = {"bd_start_c": T_I, "bd_start_r": T_I, "n0": T_I, "n0m1": T_I, "n1": T_I, "n1m1": T_I,
                     "out": arr(T_F), "shp": T_T, "u": arr(T_F)}
            args = [(v, types[v]) for v in sorted(used)]
            sites.append(site("hydk_jacobi_step_%s_%s" % (rk, ck), args, lp))
    return sites


def wrap_stencil():
    """stencil wrap (fn a => 0.2 *. (a (-1) 0 +. a 1 0) +. a 0 (-1) *. a 0 1 -. a 0 0 /. 3.0) u:
    the generic flat loop over shape_size with i0 = i / n1, i1 = i - i0 * n1
    and genBnd2D's BWrap arithmetic per access."""
    accesses = [((-1, 0), "w_n"), ((1, 0), "w_s"), ((0, -1), "w_w"), ((0, 1), "w_e"), ((0, 0), "w_c")]
    body = [
        assign("n0", C("RProj", 0, atomize("shp"))), assign("n1", C("RProj", 1, atomize("shp"))),
        assign("i0", bop("CDiv", "i", "n1")), assign("itmp", bop("CMul", "i0", "n1")),
        assign("i1", bop("CSub", "i", "itmp")),
    ]
    for k, ((d0, d1), res) in enumerate(accesses):
        for dim, (ia, na, d) in enumerate((("i0", "n0", d0), ("i1", "n1", d1))):
            raw, q, tmp, bi, lt0 = ("raw%d_%d" % (dim, k), "q%d_%d" % (dim, k), "tmp%d_%d" % (dim, k),
                                    "bi%d_%d" % (dim, k), "lt%d_%d" % (dim, k))
            body += [assign(raw, bop("CAdd", ia, d)), assign(q, bop("CDiv", raw, na)),
                     assign(tmp, bop("CMul", q, na)), assign(bi, bop("CSub", raw, tmp)),
                     assign(lt0, bop("CLt", bi, 0)),
                     sif(lt0, [assign(bi, bop("CAdd", bi, na))])]
        body += [assign("n1bi0_%d" % k, bop("CMul", "bi0_%d" % k, "n1")),
                 assign("fbi_%d" % k, bop("CAdd", "n1bi0_%d" % k, "bi1_%d" % k)),
                 assign(res, load("u", "fbi_%d" % k))]
    body += [assign("s1", bop("CAddF", "w_n", "w_s")), assign("s2", bop("CMulF", 0.2, "s1")),
             assign("s3", bop("CMulF", "w_w", "w_e")), assign("s4", bop("CAddF", "s2", "s3")),
             assign("s5", bop("CDivF", "w_c", 3.0)), assign("s6", bop("CSubF", "s4", "s5")),
             write("out", "i", "s6")]
    lp = loop(["i"], ["n"], body, par(), role="LoopPlain")
    return [site("hydk_wrap_stencil_0", [("n", T_I), ("out", arr(T_F)), ("shp", T_T), ("u", arr(T_F))], lp)]


def mirror_1d():
    """stencil mirror (fn a => 0.5 *. a (-1) +. a 0 -. 0.25 *. a 2) v, rank 1:
    genBnd1D with BMirror (period 2 (n - 1), reflection past n - 1)."""
    accesses = [(-1, "m_l"), (0, "m_c"), (2, "m_r")]
    body = []
    for k, (d, res) in enumerate(accesses):
        raw, nm1, period, q, tmp, r, lt0, geN, bi = ("%s_%d" % (x, k) for x in
                                                   ("raw", "nm1", "period", "q", "tmp", "r", "lt0", "geN", "bi"))
        body += [assign(raw, bop("CAdd", "i", d)), assign(nm1, bop("CSub", "n", 1)),
                 assign(period, bop("CMul", 2, nm1)), assign(q, bop("CDiv", raw, period)),
                 assign(tmp, bop("CMul", q, period)), assign(r, bop("CSub", raw, tmp)),
                 assign(lt0, bop("CLt", r, 0)), sif(lt0, [assign(r, bop("CAdd", r, period))]),
                 assign(geN, bop("CGe", r, "n")), sif(geN, [assign(r, bop("CSub", period, r))]),
                 assign(bi, atom(r)), assign(res, load("v", bi))]
    body += [assign("p1", bop("CMulF", 0.5, "m_l")), assign("p2", bop("CAddF", "p1", "m_c")),
             assign("p3", bop("CMulF", 0.25, "m_r")), assign("p4", bop("CSubF", "p2", "p3")),
             write("out", "i", "p4")]
    lp = loop(["i"], ["n"], body, par(), role="LoopPlain")
    return [site("hydk_mirror_1d_0", [("n", T_I), ("out", arr(T_F)), ("v", arr(T_F))], lp)]


def histogram():
    body = [assign("x", load("data", "i")), assign("idx", bop("CMod", "x", "nb")), assign("val", bop("CMul", "x", "x")),
            assign("old", load("hist", "idx")), assign("new", bop("CAdd", "val", "old")), write("hist", "idx", "new")]
    lp = loop(["i"], ["n"], body, par("ParallelScatterPrivatizedIntAdd"), role="LoopPlain")
    return [site("hydk_histogram_0", [("data", arr(T_I)), ("hist", arr(T_I)), ("n", T_I), ("nb", T_I)], lp)]


def splat_f64():
    """scatter_guarded (+.) dst (generate [n] route) (generate [n] value)
    (generate [n] keep), route i = (i * 17 + 3) % nvox, value i =
    float_of (i % 13) *. 0.125 +. 0.1, keep i = i % 3 == 0."""
    body = [
        assign("r1", bop("CMul", "i", 17)), assign("r2", bop("CAdd", "r1", 3)), assign("idx", bop("CMod", "r2", "nvox")),
        assign("v1", bop("CMod", "i", 13)), assign("v2", uop("CFloatOf", "v1")), assign("v3", bop("CMulF", "v2", 0.125)),
        assign("val", bop("CAddF", "v3", 0.1)),
        assign("g1", bop("CMod", "i", 3)), assign("guard", bop("CEq", "g1", 0)),
        sif("guard", [assign("old", load("dst", "idx")), assign("new", bop("CAddF", "val", "old")),
                      write("dst", "idx", "new")]),
    ]
    lp = loop(["i"], ["n"], body, par("ParallelScatterPrivatizedFloatAdd"), role="LoopPlain")
    return [site("hydk_splat_f64_0", [("dst", arr(T_F)), ("n", T_I), ("nvox", T_I)], lp)]


def row_sums():
    inner = loop(["k"], ["red_dim"], [
        assign("flat_in", bop("CAdd", "base", "k")), assign("elem", load("m", "flat_in")),
        assign("acc", bop("CAddF", "acc", "elem")),
    ], SERIAL, red=("acc", 0, "RAdd"), role="LoopReduction")
    body = [assign("base", bop("CMul", "j", "red_dim")), assign("acc", atom(0.0)), inner, write("out", "j", "acc")]
    lp = loop(["j"], ["out_n"], body, par(), role="LoopMapReduction")
    return [site("hydk_row_sums_0", [("m", arr(T_F)), ("out", arr(T_F)), ("out_n", T_I), ("red_dim", T_I)], lp)]


def gather():
    body = [assign("p", load("perm", "i")), assign("v", load("src", "p")), write("out", "i", "v")]
    lp = loop(["i"], ["n"], body, par(), role="LoopMap")
    return [site("hydk_gather_0", [("n", T_I), ("out", arr(T_I)), ("perm", arr(T_I)), ("src", arr(T_I))], lp)]


def map3d():
    body = [
        assign("nd", C("RTuple", [atomize("a"), atomize("b"), atomize("c")])),
        assign("flat", C("RNdToFlat", atomize("nd"), atomize("shp"))),
        assign("fa", uop("CFloatOf", "a")), assign("fb", uop("CFloatOf", "b")), assign("fc", uop("CFloatOf", "c")),
        assign("x", load("src", "flat")), assign("m1", bop("CMulF", 2.0, "fb")), assign("m2", bop("CSubF", "fa", "m1")),
        assign("m3", bop("CMulF", 0.5, "fc")), assign("m4", bop("CAddF", "m2", "m3")), assign("m5", bop("CAddF", "m4", "x")),
        write("out", "flat", "m5"),
    ]
    lp = loop(["a", "b", "c"], ["n0", "n1", "n2"], body, par(), role="LoopMap")
    return [site("hydk_map3d_0", [("n0", T_I), ("n1", T_I), ("n2", T_I), ("out", arr(T_F)), ("shp", T_T),
                                  ("src", arr(T_F))], lp)]


def int_iterate():
    """iterate steps (fn u => stencil clamp f u) u0 on an int64 grid, f = the
    5-point sum with the centre tripled, C % 1000003: the split step's 9 region
    loops (jacobi_step's shape with this body) under the large-grid branch; no
    small-grid loop (grids below 2 x 2 run no loop and keep the fresh array's
    zeros)."""
    regions = [s.get("hsLoop") for s in jacobi_step(int_body)]
    step = [
        assign("shp", C("RArrayShape", atomize("u"))),
        assign("out", C("RArrayAlloc", atomize("shp"))),
        assign("n0", C("RProj", 0, atomize("shp"))),
        assign("n1", C("RProj", 1, atomize("shp"))),
        assign("n0m1", bop("CSub", "n0", 1)), assign("n1m1", bop("CSub", "n1", 1)),
        assign("bd_start_r", bop("CSub", "n0", 1)), assign("bd_start_c", bop("CSub", "n1", 1)),
        assign("rows_large", bop("CGe", "n0", 2)), assign("cols_large", bop("CGe", "n1", 2)),
        assign("large", bop("CAnd", "rows_large", "cols_large")),
        sif("large", regions, []),
        assign("u", atom("out")),
    ]
    lp = loop(["t"], ["steps"], step, SERIAL, role="LoopIterate")
    return [site("hydk_int_iterate", [("steps", T_I), ("u", arr(T_I))], lp, result=("u", arr(T_I)),
                 extra_types=[("out", arr(T_I))])]


def heat_iterate():
    """iterate steps (fn u => stencil clamp heat u) u0: Lowering.hs:840-857 lowers
    the iterate to a serial LoopIterate loop whose step is the stencil's statements
    followed by u = out.  The step: the input's shape, the new array (RArrayAlloc),
    the extents and clamp bounds of lowerSplitStencil2D (:2562-2600), then a branch
    on both extents being at least the footprint (2): the 9 region loops of
    jacobi_step, or for smaller grids the generic flat loop over shape_size
    (:1997-2030) with genBnd1DVar's BClamp index arithmetic per access
    (:2760-2770)."""
    regions = [s.get("hsLoop") for s in jacobi_step()]
    small = [assign("i0", bop("CDiv", "i", "n1")), assign("itmp", bop("CMul", "i0", "n1")),
             assign("i1", bop("CSub", "i", "itmp"))]
    for (dr, dc), res in HEAT_ACCESSES:
        comps = []
        for dim, (ia, na, d) in enumerate((("i0", "n0", dr), ("i1", "n1", dc))):
            raw, bi, lt0, nm1, gen = ("s%s_%d_%s" % (x, dim, res) for x in ("raw", "bi", "lt0", "nm1", "geN"))
            small += [assign(raw, bop("CAdd", ia, d)), assign(bi, atom(raw)), assign(lt0, bop("CLt", raw, 0)),
                      sif(lt0, [assign(bi, atom(0))]), assign(nm1, bop("CSub", na, 1)),
                      assign(gen, bop("CGe", bi, na)), sif(gen, [assign(bi, atom(nm1))])]
            comps.append(bi)
        small += [assign("sidx_" + res, C("RTuple", [atomize(comps[0]), atomize(comps[1])])),
                  assign("sflat_" + res, C("RNdToFlat", atomize("sidx_" + res), atomize("shp"))),
                  assign(res, load("u", "sflat_" + res))]
    small += heat_body("val") + [write("out", "i", "val")]
    step = [
        assign("shp", C("RArrayShape", atomize("u"))),
        assign("out", C("RArrayAlloc", atomize("shp"))),
        assign("n", C("RShapeSize", atomize("shp"))),
        assign("n0", C("RProj", 0, atomize("shp"))),
        assign("n1", C("RProj", 1, atomize("shp"))),
        assign("n0m1", bop("CSub", "n0", 1)),
        assign("n1m1", bop("CSub", "n1", 1)),
        assign("bd_start_r", bop("CSub", "n0", 1)),
        assign("bd_start_c", bop("CSub", "n1", 1)),
        assign("rows_large", bop("CGe", "n0", 2)),
        assign("cols_large", bop("CGe", "n1", 2)),
        assign("large", bop("CAnd", "rows_large", "cols_large")),
        sif("large", regions, [loop(["i"], ["n"], small, par(), role="LoopMap")]),
        assign("u", atom("out")),
    ]
    lp = loop(["t"], ["steps"], step, SERIAL, role="LoopIterate")
    return [site("hydk_heat_iterate", [("steps", T_I), ("u", arr(T_F))], lp, result=("u", arr(T_F)),
                 extra_types=[("out", arr(T_F))])]


# The 4th-order heat body on a 9-point cross of radius 2:
# a 0 0 +. 0.01/12 *. (16 *. (a(-1)0 +. a 1 0 +. a 0(-1) +. a 0 1) -. (a(-2)0 +. a 2 0 +. a 0(-2) +. a 0 2)
#                      -. 60 *. a 0 0)
WIDE_ACCESSES = [((-1, 0), "b_n1"), ((1, 0), "b_s1"), ((0, -1), "b_w1"), ((0, 1), "b_e1"), ((-2, 0), "b_n2"),
                 ((2, 0), "b_s2"), ((0, -2), "b_w2"), ((0, 2), "b_e2"), ((0, 0), "b_c")]
WIDE_COEF = 0.01 / 12.0


def wide_body(res):
    return [
        assign("q1", bop("CAddF", "b_n1", "b_s1")), assign("q2", bop("CAddF", "b_w1", "b_e1")),
        assign("q3", bop("CAddF", "q1", "q2")), assign("q4", bop("CAddF", "b_n2", "b_s2")),
        assign("q5", bop("CAddF", "b_w2", "b_e2")), assign("q6", bop("CAddF", "q4", "q5")),
        assign("q7", bop("CMulF", 16.0, "q3")), assign("q8", bop("CSubF", "q7", "q6")),
        assign("q9", bop("CMulF", 60.0, "b_c")), assign("q10", bop("CSubF", "q8", "q9")),
        assign("q11", bop("CMulF", WIDE_COEF, "q10")), assign(res, bop("CAddF", "b_c", "q11")),
    ]


def wide_iterate():
    """iterate steps (fn u => stencil clamp heat4 u) u0 with a footprint of
    radius 2: lowerSplitStencil2D (:2562-2665) splits rows and columns into a
    2-row/column left edge (one segment per row, offsets 0 and 1), the interior
    (offset 2, n - 4 iterations) and a 2-row/column right edge (offsets n - 2,
    n - 1): 25 region loops, each access's clamp folded per segment as
    specializeClampIndex1D does (an index past the edge becomes the edge,
    any other stays iterator + offset); grids below 4 x 4 take the generic flat
    loop with genBnd1DVar's BClamp arithmetic.  The step ends with u = out
    (Lowering.hs:840-857)."""
    segs0 = [("l0", 1, 0), ("l1", 1, 1), ("in", C("IAdd", C("IVar", "n0"), C("IConst", -4)), 2),
             ("r0", 1, "bd_r0"), ("r1", 1, "bd_r1")]
    segs1 = [("l0", 1, 0), ("l1", 1, 1), ("in", C("IAdd", C("IVar", "n1"), C("IConst", -4)), 2),
             ("r0", 1, "bd_c0"), ("r1", 1, "bd_c1")]
    regions = []
    for rk, rbound, roff in segs0:
        for ck, cbound, coff in segs1:
            body = []

            def shift(it, off, tag):
                if off == 0:
                    return it
                v = "%s_%s" % (it, tag)
                body.append(assign(v, bop("CAdd", it, off)))
                return v

            ra = shift("i", roff, "r")
            ca = shift("j", coff, "c")

            def fold(kind, base, off, nm1, tag):
                # segment l0 / l1 is row (column) 0 / 1, r0 / r1 is n - 2 /
                # n - 1: an access past the edge folds to the edge
                if kind in ("l0", "l1") and int(kind[1]) + off < 0:
                    return 0
                if kind in ("r0", "r1") and off > 1 - int(kind[1]):
                    return nm1
                if off == 0:
                    return base
                v = "%s_%s" % (base, tag)
                body.append(assign(v, bop("CAdd", base, off)))
                return v

            for (dr, dc), res in WIDE_ACCESSES:
                tag = "%s%d%s%d" % ("m" if dr < 0 else "p", abs(dr), "m" if dc < 0 else "p", abs(dc))
                r = fold(rk, ra, dr, "n0m1", "r" + tag)
                c = fold(ck, ca, dc, "n1m1", "c" + tag)
                body.append(assign("idx_" + res, C("RTuple", [atomize(r), atomize(c)])))
                body.append(assign("flat_" + res, C("RNdToFlat", atomize("idx_" + res), atomize("shp"))))
                body.append(assign(res, load("u", "flat_" + res)))
            body += wide_body("val")
            body.append(assign("idx_o", C("RTuple", [atomize(ra), atomize(ca)])))
            body.append(assign("flat_o", C("RNdToFlat", atomize("idx_o"), atomize("shp"))))
            body.append(write("out", "flat_o", "val"))
            regions.append(loop(["i", "j"], [rbound, cbound], body, par(), role="LoopMap"))
    small = [assign("i0", bop("CDiv", "i", "n1")), assign("itmp", bop("CMul", "i0", "n1")),
             assign("i1", bop("CSub", "i", "itmp"))]
    for (dr, dc), res in WIDE_ACCESSES:
        comps = []
        for dim, (ia, na, d) in enumerate((("i0", "n0", dr), ("i1", "n1", dc))):
            raw, bi, lt0, nm1, gen = ("w%s_%d_%s" % (x, dim, res) for x in ("raw", "bi", "lt0", "nm1", "geN"))
            small += [assign(raw, bop("CAdd", ia, d)), assign(bi, atom(raw)), assign(lt0, bop("CLt", raw, 0)),
                      sif(lt0, [assign(bi, atom(0))]), assign(nm1, bop("CSub", na, 1)),
                      assign(gen, bop("CGe", bi, na)), sif(gen, [assign(bi, atom(nm1))])]
            comps.append(bi)
        small += [assign("widx_" + res, C("RTuple", [atomize(comps[0]), atomize(comps[1])])),
                  assign("wflat_" + res, C("RNdToFlat", atomize("widx_" + res), atomize("shp"))),
                  assign(res, load("u", "wflat_" + res))]
    small += wide_body("val") + [write("out", "i", "val")]
    step = [
        assign("shp", C("RArrayShape", atomize("u"))),
        assign("out", C("RArrayAlloc", atomize("shp"))),
        assign("n", C("RShapeSize", atomize("shp"))),
        assign("n0", C("RProj", 0, atomize("shp"))),
        assign("n1", C("RProj", 1, atomize("shp"))),
        assign("n0m1", bop("CSub", "n0", 1)), assign("n1m1", bop("CSub", "n1", 1)),
        assign("bd_r0", bop("CSub", "n0", 2)), assign("bd_r1", bop("CSub", "n0", 1)),
        assign("bd_c0", bop("CSub", "n1", 2)), assign("bd_c1", bop("CSub", "n1", 1)),
        assign("rows_large", bop("CGe", "n0", 4)), assign("cols_large", bop("CGe", "n1", 4)),
        assign("large", bop("CAnd", "rows_large", "cols_large")),
        sif("large", regions, [loop(["i"], ["n"], small, par(), role="LoopMap")]),
        assign("u", atom("out")),
    ]
    lp = loop(["t"], ["steps"], step, SERIAL, role="LoopIterate")
    return [site("hydk_wide_iterate", [("steps", T_I), ("u", arr(T_F))], lp, result=("u", arr(T_F)),
                 extra_types=[("out", arr(T_F))])]


def interior_iterate():
    """iterate steps (fn u => a heat step that writes only the interior of a
    fresh array) u0: the step's one loop writes out[(i + 1, j + 1)] for the
    (n0 - 2) x (n1 - 2) interior, so each step's boundary keeps the zero the
    new array is allocated with (calloc, runtime/hydrangea_runtime.h:292-310).
    Its loop qualifies for the temporal form when hydgen plans it, but its box
    does not tile the grid: the launcher must find that at run time and run
    the steps in order, each into a zeroed array."""
    body = [assign("i_r", bop("CAdd", "i", 1)), assign("j_c", bop("CAdd", "j", 1))]
    for (dr, dc), res in HEAT_ACCESSES:
        r = "i_r" if dr == 0 else "ir_%s" % res
        c = "j_c" if dc == 0 else "jc_%s" % res
        if dr:
            body.append(assign(r, bop("CAdd", "i_r", dr)))
        if dc:
            body.append(assign(c, bop("CAdd", "j_c", dc)))
        body += [assign("idx_" + res, C("RTuple", [atomize(r), atomize(c)])),
                 assign("flat_" + res, C("RNdToFlat", atomize("idx_" + res), atomize("shp"))),
                 assign(res, load("u", "flat_" + res))]
    body += heat_body("val")
    body += [assign("idx_o", C("RTuple", [atomize("i_r"), atomize("j_c")])),
             assign("flat_o", C("RNdToFlat", atomize("idx_o"), atomize("shp"))), write("out", "flat_o", "val")]
    inner = loop(["i", "j"], ["rows", "cols"], body, par(), role="LoopMap")
    step = [
        assign("shp", C("RArrayShape", atomize("u"))),
        assign("out", C("RArrayAlloc", atomize("shp"))),
        assign("n0", C("RProj", 0, atomize("shp"))),
        assign("n1", C("RProj", 1, atomize("shp"))),
        assign("rows", bop("CSub", "n0", 2)), assign("cols", bop("CSub", "n1", 2)),
        inner,
        assign("u", atom("out")),
    ]
    lp = loop(["t"], ["steps"], step, SERIAL, role="LoopIterate")
    return [site("hydk_interior_iterate", [("steps", T_I), ("u", arr(T_F))], lp, result=("u", arr(T_F)),
                 extra_types=[("out", arr(T_F))])]


PROGRAMS = {
    "map_chain": map_chain, "map_chain2": map_chain2, "dot": dot, "reduce_i64": reduce_i64,
    "jacobi_step": jacobi_step, "wrap_stencil": wrap_stencil, "mirror_1d": mirror_1d, "histogram": histogram,
    "splat_f64": splat_f64, "row_sums": row_sums, "gather": gather, "map3d": map3d,
    "heat_iterate": heat_iterate, "wide_iterate": wide_iterate, "interior_iterate": interior_iterate,
    "int_iterate": int_iterate,
}


def group_overlap():
    """Three maps the printed C would call back to back: out = xs + 1 over
    [0, n), out[i + sh] = 2 xs[i + sh] over [0, n - sh) (its writes overlap
    the first's), and ys = out + xs (it reads what the others write).  Their
    group launcher must find them dependent and run them in order."""
    a = loop(["i"], ["n"], [assign("x", load("xs", "i")), assign("y", bop("CAddF", "x", 1.0)), write("out", "i", "y")],
             par())
    b = loop(["i"], [C("ISub", C("IVar", "n"), C("IVar", "sh"))],
             [assign("ii", bop("CAdd", "i", "sh")), assign("x", load("xs", "ii")), assign("y", bop("CMulF", "x", 2.0)),
              write("out", "ii", "y")], par())
    c = loop(["i"], ["n"], [assign("o", load("out", "i")), assign("x", load("xs", "i")), assign("y", bop("CAddF", "o", "x")),
                            write("ys", "i", "y")], par())
    return [site("hydk_overlap_add1", [("n", T_I), ("out", arr(T_F)), ("xs", arr(T_F))], a),
            site("hydk_overlap_twice", [("n", T_I), ("out", arr(T_F)), ("sh", T_I), ("xs", arr(T_F))], b),
            site("hydk_overlap_sum", [("n", T_I), ("out", arr(T_F)), ("xs", arr(T_F)), ("ys", arr(T_F))], c)]


PROGRAMS["group_overlap"] = group_overlap

# Runs of sites the printed C calls through one group launcher (HipGroup
# lines): the split stencil's 9 region loops, which follow one another with
# nothing between them, and the overlap program's three maps.
GROUPS = {
    "jacobi_step": [("hydk_jacobi_step", None)],
    "group_overlap": [("hydk_overlap", None)],
}


def groups(name):
    """HipGroup records of a program: each group takes every site of the
    program, in order."""
    out = []
    for gname, members in GROUPS.get(name, []):
        ms = members or [s.get("hsName") for s in sites(name)]
        out.append(R("HipGroup", [("hgName", gname), ("hgSites", list(ms))]))
    return out


def sites(name):
    return PROGRAMS[name]()

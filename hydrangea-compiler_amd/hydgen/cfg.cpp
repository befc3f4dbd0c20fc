// cfg.cpp - derived-Show values -> the C++ mirror of the reference's CFG.
//
// Constructors and fields follow src/Language/Hydrangea/CFGCore.hs (BinOp
// :33-57, UnOp :60-74, Atom :77-85, CType :103-113, RHS :140-169, Redop :236)
// and src/Language/Hydrangea/CFG.hs (IndexExpr :55-79, ParallelStrategy
// :184-191, ParallelSpec :199-204, ReductionSpec :207-212, LoopRole :226-236,
// ExecPolicy :239-244, LoopSpec :251-259, Stmt :262-270).  A HipSite record
// (Language.Hydrangea.CodegenHIP) wraps one loop with its arguments and types.
#include <errno.h>
#include <math.h>
#include <stdlib.h>

#include "hydgen.hpp"

namespace hydgen {

namespace {

[[noreturn]] void bad(const Value& v, const std::string& what) {
    throw Error("CFG: " + what + " (got " + (v.text.empty() ? std::string("value") : v.text) + ")");
}

const Value& arg(const Value& v, size_t k) {
    if (v.kind != Value::Con || v.items.size() <= k) bad(v, "missing constructor argument");
    return *v.items[k];
}

std::string str(const Value& v) {
    if (v.kind != Value::Str) bad(v, "expected a string");
    return v.text;
}

int64_t integer(const Value& v) {
    if (v.kind != Value::Num) bad(v, "expected an integer");
    errno = 0;
    char* end = nullptr;
    long long x = strtoll(v.text.c_str(), &end, 10);
    if (errno || !end || *end) bad(v, "integer out of int64 range or malformed");
    return (int64_t)x;
}

double floating(const Value& v) {
    if (v.kind != Value::Num) bad(v, "expected a double");
    if (v.text == "Infinity") return HUGE_VAL;
    if (v.text == "-Infinity") return -HUGE_VAL;
    if (v.text == "NaN") return NAN;
    char* end = nullptr;
    double x = strtod(v.text.c_str(), &end);  // shortest round-trip digits: the same double
    if (!end || *end) bad(v, "malformed double");
    return x;
}

const std::vector<VPtr>& list(const Value& v) {
    if (v.kind != Value::List) bad(v, "expected a list");
    return v.items;
}

BinOp binop(const Value& v) {
    static const char* names[] = {"CAdd", "CSub", "CMul", "CDiv", "CMod", "CEq", "CNeq", "CLt",
                                  "CLe", "CGt", "CGe", "CAnd", "COr", "CAddF", "CSubF", "CMulF",
                                  "CDivF", "CEqF", "CNeqF", "CLtF", "CLeF", "CGtF", "CGeF"};
    for (int k = 0; k < (int)(sizeof names / sizeof *names); ++k)
        if (v.is_con(names[k])) return (BinOp)k;
    bad(v, "unknown BinOp");
}

UnOp unop(const Value& v) {
    static const char* names[] = {"CNot", "CNeg", "CSqrt", "CExpF", "CLog", "CSin", "CCos",
                                  "CAbsF", "CFloorF", "CCeilF", "CErf", "CFloatOf", "CIntOf"};
    for (int k = 0; k < (int)(sizeof names / sizeof *names); ++k)
        if (v.is_con(names[k])) return (UnOp)k;
    bad(v, "unknown UnOp");
}

Atom atom(const Value& v) {
    Atom a;
    if (v.is_con("AVar")) {
        a.kind = Atom::Var;
        a.var = str(arg(v, 0));
    } else if (v.is_con("AInt")) {
        a.kind = Atom::Int;
        a.i = integer(arg(v, 0));
    } else if (v.is_con("AFloat")) {
        a.kind = Atom::Float;
        a.f = floating(arg(v, 0));
    } else if (v.is_con("ABool")) {
        a.kind = Atom::Bool;
        a.b = arg(v, 0).is_con("True");
    } else if (v.is_con("AUnit")) {
        a.kind = Atom::Unit;
    } else if (v.is_con("AString")) {
        a.kind = Atom::String;
        a.var = str(arg(v, 0));
    } else if (v.is_con("AVecVar")) {
        a.kind = Atom::VecVar;
        a.var = str(arg(v, 0));
    } else {
        bad(v, "unknown Atom");
    }
    return a;
}

CType ctype(const Value& v) {
    CType t;
    if (v.is_con("CTInt64")) t.kind = CType::Int64;
    else if (v.is_con("CTDouble")) t.kind = CType::Double;
    else if (v.is_con("CTBool")) t.kind = CType::Bool;
    else if (v.is_con("CTUnit")) t.kind = CType::Unit;
    else if (v.is_con("CTTuple")) t.kind = CType::Tuple;
    else if (v.is_con("CTArray")) {
        t.kind = CType::Array;
        t.elem = std::make_shared<CType>(ctype(arg(v, 0)));
    } else if (v.is_con("CTPair")) t.kind = CType::Pair;
    else if (v.is_con("CTRecord")) t.kind = CType::Record;
    else if (v.is_con("CTUnknown")) t.kind = CType::Unknown;
    else bad(v, "unknown CType");
    return t;
}

RHS rhs(const Value& v) {
    struct Shape {
        const char* name;
        RHS::Kind kind;
        int natoms;  // leading Atom arguments (after any non-atom ones handled below)
    };
    static const Shape shapes[] = {
        {"RAtom", RHS::RAtom, 1},           {"RArrayAlloc", RHS::RArrayAlloc, 1},
        {"RArrayCopy", RHS::RArrayCopy, 1}, {"RArrayLoad", RHS::RArrayLoad, 2},
        {"RArrayShape", RHS::RArrayShape, 1}, {"RShapeSize", RHS::RShapeSize, 1},
        {"RShapeInit", RHS::RShapeInit, 1}, {"RShapeLast", RHS::RShapeLast, 1},
        {"RFlatToNd", RHS::RFlatToNd, 2},   {"RNdToFlat", RHS::RNdToFlat, 2},
        {"R2DToFlat", RHS::R2DToFlat, 2},   {"RVecLoad", RHS::RVecLoad, 2},
        {"RVecStore", RHS::RVecStore, 3},   {"RVecSplat", RHS::RVecSplat, 1},
        {"RArrayFree", RHS::RArrayFree, 1},
    };
    RHS r;
    for (const Shape& sh : shapes)
        if (v.is_con(sh.name)) {
            r.kind = sh.kind;
            for (int k = 0; k < sh.natoms; ++k) r.args.push_back(atom(arg(v, k)));
            return r;
        }
    if (v.is_con("RBinOp") || v.is_con("RVecBinOp")) {
        r.kind = v.is_con("RBinOp") ? RHS::RBinOp : RHS::RVecBinOp;
        r.bop = binop(arg(v, 0));
        r.args = {atom(arg(v, 1)), atom(arg(v, 2))};
    } else if (v.is_con("RUnOp") || v.is_con("RVecUnOp")) {
        r.kind = v.is_con("RUnOp") ? RHS::RUnOp : RHS::RVecUnOp;
        r.uop = unop(arg(v, 0));
        r.args = {atom(arg(v, 1))};
    } else if (v.is_con("RTuple")) {
        r.kind = RHS::RTuple;
        for (const VPtr& a : list(arg(v, 0))) r.args.push_back(atom(*a));
    } else if (v.is_con("RProj")) {
        r.kind = RHS::RProj;
        r.proj = integer(arg(v, 0));
        r.args = {atom(arg(v, 1))};
    } else if (v.is_con("RCall")) {
        r.kind = RHS::RCall;
        r.name = str(arg(v, 0));
        for (const VPtr& a : list(arg(v, 1))) r.args.push_back(atom(*a));
    } else if (v.is_con("RVecReduce")) {
        r.kind = RHS::RVecReduce;
        r.bop = binop(arg(v, 0));
        r.args = {atom(arg(v, 1))};
    } else if (v.is_con("RRecord") || v.is_con("RRecordProj") || v.is_con("RPairMake") ||
               v.is_con("RPairFst") || v.is_con("RPairSnd")) {
        // records and pairs: kept as a kind only (never offloaded)
        r.kind = v.is_con("RRecord")       ? RHS::RRecord
                 : v.is_con("RRecordProj") ? RHS::RRecordProj
                 : v.is_con("RPairMake")   ? RHS::RPairMake
                 : v.is_con("RPairFst")    ? RHS::RPairFst
                                           : RHS::RPairSnd;
    } else {
        bad(v, "unknown RHS");
    }
    return r;
}

IndexExpr index_expr(const Value& v) {
    IndexExpr e;
    if (v.is_con("IVar")) {
        e.kind = IndexExpr::IVar;
        e.var = str(arg(v, 0));
    } else if (v.is_con("IConst")) {
        e.kind = IndexExpr::IConst;
        e.c = integer(arg(v, 0));
    } else if (v.is_con("IAdd") || v.is_con("ISub") || v.is_con("IMul") || v.is_con("IDiv") ||
               v.is_con("IFlatToNd") || v.is_con("INdToFlat")) {
        e.kind = v.is_con("IAdd")        ? IndexExpr::IAdd
                 : v.is_con("ISub")      ? IndexExpr::ISub
                 : v.is_con("IMul")      ? IndexExpr::IMul
                 : v.is_con("IDiv")      ? IndexExpr::IDiv
                 : v.is_con("IFlatToNd") ? IndexExpr::IFlatToNd
                                         : IndexExpr::INdToFlat;
        e.kids = {index_expr(arg(v, 0)), index_expr(arg(v, 1))};
    } else if (v.is_con("ITuple")) {
        e.kind = IndexExpr::ITuple;
        for (const VPtr& k : list(arg(v, 0))) e.kids.push_back(index_expr(*k));
    } else if (v.is_con("IProj")) {
        e.kind = IndexExpr::IProj;
        e.c = integer(arg(v, 0));
        e.kids = {index_expr(arg(v, 1))};
    } else if (v.is_con("ICall")) {
        e.kind = IndexExpr::ICall;
        e.var = str(arg(v, 0));
        for (const VPtr& k : list(arg(v, 1))) e.kids.push_back(index_expr(*k));
    } else {
        bad(v, "unknown IndexExpr");
    }
    return e;
}

Strategy strategy(const Value& v) {
    if (v.is_con("ParallelGeneric")) return Strategy::Generic;
    if (v.is_con("ParallelScatterDirect")) return Strategy::ScatterDirect;
    if (v.is_con("ParallelScatterAtomicAddInt")) return Strategy::ScatterAtomicAddInt;
    if (v.is_con("ParallelScatterAtomicAddFloat")) return Strategy::ScatterAtomicAddFloat;
    if (v.is_con("ParallelScatterPrivatizedIntAdd")) return Strategy::ScatterPrivatizedIntAdd;
    if (v.is_con("ParallelScatterPrivatizedFloatAdd")) return Strategy::ScatterPrivatizedFloatAdd;
    bad(v, "unknown ParallelStrategy");
}

Role role(const Value& v) {
    static const char* names[] = {"LoopPlain", "LoopFold", "LoopMap", "LoopReductionWrapper",
                                  "LoopReduction", "LoopMapReduction", "LoopIterate", "LoopSegRedOuter"};
    for (int k = 0; k < 8; ++k)
        if (v.is_con(names[k])) return (Role)k;
    bad(v, "unknown LoopRole");
}

LoopSpec loop_spec(const Value& v) {
    if (!v.is_con("LoopSpec") || v.kind != Value::Record) bad(v, "expected a LoopSpec record");
    LoopSpec s;
    for (const VPtr& i : list(v.field("lsIters"))) s.iters.push_back(str(*i));
    for (const VPtr& b : list(v.field("lsBounds"))) s.bounds.push_back(index_expr(*b));
    const Value& ex = v.field("lsExec");
    if (ex.is_con("Serial")) {
        s.exec = LoopSpec::Serial;
    } else if (ex.is_con("Vector")) {
        s.exec = LoopSpec::Vector;
    } else if (ex.is_con("Parallel") || ex.is_con("Workshare")) {
        s.exec = ex.is_con("Parallel") ? LoopSpec::Parallel : LoopSpec::Workshare;
        s.strategy = strategy(arg(ex, 0).field("psStrategy"));
    } else {
        bad(ex, "unknown ExecPolicy");
    }
    const Value& red = v.field("lsRed");
    if (red.is_con("Just")) {
        const Value& rs = arg(red, 0);
        s.has_red = true;
        s.red_acc = str(rs.field("rsAccVar"));
        s.red_init = index_expr(rs.field("rsInit"));
        s.red_mul = rs.field("rsRedop").is_con("RMul");
        if (!s.red_mul && !rs.field("rsRedop").is_con("RAdd")) bad(rs, "unknown Redop");
    } else if (!red.is_con("Nothing")) {
        bad(red, "lsRed: expected Maybe ReductionSpec");
    }
    s.role = role(v.field("lsRole"));
    for (const VPtr& o : list(v.field("lsOrigins"))) s.origins.push_back(index_expr(*o));
    // CFG.hs:246-249: origins default to all IConst 0 (an empty list in mkProc-built loops)
    while (s.origins.size() < s.iters.size()) s.origins.push_back(IndexExpr());
    if (s.bounds.size() != s.iters.size()) bad(v, "lsBounds and lsIters differ in length");
    return s;
}

std::vector<Stmt> stmts(const Value& v);

Stmt stmt(const Value& v) {
    Stmt s;
    if (v.is_con("SAssign")) {
        s.kind = Stmt::SAssign;
        s.var = str(arg(v, 0));
        s.rhs = rhs(arg(v, 1));
    } else if (v.is_con("SArrayWrite")) {
        s.kind = Stmt::SArrayWrite;
        s.a0 = atom(arg(v, 0));
        s.a1 = atom(arg(v, 1));
        s.a2 = atom(arg(v, 2));
    } else if (v.is_con("SLoop")) {
        s.kind = Stmt::SLoop;
        s.loop = loop_spec(arg(v, 0));
        s.body = stmts(arg(v, 1));
    } else if (v.is_con("SParallelRegion")) {
        s.kind = Stmt::SParallelRegion;
        s.body = stmts(arg(v, 0));
    } else if (v.is_con("SIf")) {
        s.kind = Stmt::SIf;
        s.a0 = atom(arg(v, 0));
        s.body = stmts(arg(v, 1));
        s.els = stmts(arg(v, 2));
    } else if (v.is_con("SReturn")) {
        s.kind = Stmt::SReturn;
        s.a0 = atom(arg(v, 0));
    } else if (v.is_con("SBreak")) {
        s.kind = Stmt::SBreak;
    } else {
        bad(v, "unknown Stmt");
    }
    return s;
}

std::vector<Stmt> stmts(const Value& v) {
    std::vector<Stmt> out;
    for (const VPtr& s : list(v)) out.push_back(stmt(*s));
    return out;
}

std::pair<std::string, CType> typed_var(const Value& v) {
    if (v.kind != Value::Tuple || v.items.size() != 2) bad(v, "expected (name, CType)");
    return {str(*v.items[0]), ctype(*v.items[1])};
}

}  // namespace

std::string CType::show() const {
    switch (kind) {
        case Int64: return "CTInt64";
        case Double: return "CTDouble";
        case Bool: return "CTBool";
        case Unit: return "CTUnit";
        case Tuple: return "CTTuple";
        case Array: return "CTArray (" + (elem ? elem->show() : std::string("?")) + ")";
        case Pair: return "CTPair";
        case Record: return "CTRecord";
        default: return "CTUnknown";
    }
}

Site site_from_value(const Value& v) {
    if (!v.is_con("HipSite") || v.kind != Value::Record) bad(v, "expected a HipSite record");
    Site s;
    s.name = str(v.field("hsName"));
    for (const VPtr& a : list(v.field("hsArgs"))) s.args.push_back(typed_var(*a));
    const Value& res = v.field("hsResult");
    if (res.is_con("Just")) {
        s.has_result = true;
        s.result = typed_var(arg(res, 0));
    } else if (!res.is_con("Nothing")) {
        bad(res, "hsResult: expected Maybe (name, CType)");
    }
    const Value& types = v.field("hsTypes");
    if (!types.is_con("fromList")) bad(types, "hsTypes: expected a Data.Map");
    for (const VPtr& kv : list(arg(types, 0))) {
        auto tv = typed_var(*kv);
        s.types[tv.first] = tv.second;
    }
    s.loop = stmt(v.field("hsLoop"));
    if (s.loop.kind != Stmt::SLoop) bad(v, "hsLoop must be an SLoop");
    return s;
}

std::vector<Site> read_sites(const std::string& text) {
    std::vector<Site> out;
    for (const VPtr& v : parse_show_lines(text))
        if (!v->is_con("HipGroup")) out.push_back(site_from_value(*v));
    return out;
}

std::vector<Group> read_groups(const std::string& text) {
    std::vector<Group> out;
    for (const VPtr& v : parse_show_lines(text)) {
        if (!v->is_con("HipGroup")) continue;
        if (v->kind != Value::Record) bad(*v, "expected a HipGroup record");
        Group g;
        g.name = str(v->field("hgName"));
        for (const VPtr& m : list(v->field("hgSites"))) g.members.push_back(str(*m));
        out.push_back(g);
    }
    return out;
}

}  // namespace hydgen

// codegen.cpp - the GPU-schedule pass and the HIP emitter of hydgen.
//
// Schedule (per site; the site's loop is the outermost parallel loop of a
// fused nest, chosen by Language.Hydrangea.CodegenHIP):
//   * reduction  - a loop with a ReductionSpec (CFG.hs:207-212), or a
//     LoopReductionWrapper of trip count 1 around one (Lowering.hs:1150-1182):
//     a streaming grid, `unroll` independent partial accumulators per thread
//     (each starting at the operator's identity), a fixed wave64 xor-shuffle
//     tree, the waves of a workgroup in order, then one workgroup folding the
//     workgroup partials in a fixed tree and combining the accumulator's
//     value on entry: deterministic for a given device, and for fp64 a
//     reassociation of the serial sum (the reference's OpenMP
//     reduction(+:acc) reassociates too, CodegenC.hs:1306-1413);
//   * scatter    - a 1-D loop whose body holds the combine of
//     Lowering.hs:3553-3557 (old = dst[r]; new = val + old; dst[r] = new),
//     optionally under guards: int64 (+) accumulates in LDS-private bins
//     (when the destination fits 64 KiB) merged through spread copies of a
//     zeroed scratch block and a fold launch, or with global 64-bit atomics
//     beyond that - integer + is order-free, so the bits equal the serial
//     loop's; fp64 (+.) writes (route, value, taken) per element and hands
//     them to the library's order-exact scatter, so each bin receives its
//     values in element order, bit-exact to the serial loop;
//   * map, 1-D   - a streaming grid (a few workgroups per CU, grid-stride),
//     `unroll` elements per thread per round spaced a workgroup apart so each
//     wave's accesses are contiguous; when every array write is a top-level
//     statement and no written array is read, the round computes all its
//     elements first and stores afterwards (the loads of later elements
//     issue before the stores of earlier ones); loads and stores at the loop
//     index of arrays touched nowhere else are non-temporal, and when all of
//     those arrays are 16-byte aligned each thread moves two adjacent
//     elements per access (a 16-byte load or store);
//   * map, 2-D   - workgroups of 256 threads over 256 consecutive inner
//     (contiguous) indices; grid.y splits the outer index into bands of
//     consecutive rows, each walked R = 4 rows per round (computed first,
//     stored after) so a row's neighbours are read while still in L2 (the
//     clamp-stencil region loops, Lowering.hs:2159-2172);
//   * map, N-D   - flattened, the indices recovered by division.
// A map form's body is a device function of its workgroup's grid position,
// its arguments a struct the launcher's prep function fills; a group of
// adjacent sites (HipGroup) gets one kernel that runs each member's body on
// its own range of workgroups, launched when the members' accesses are
// independent at run time (no member reads bytes another writes; shared
// written arrays only through disjoint boxes of the iteration space), else
// the members' launchers run in order.
// Loop-invariant statements at the head of the body (the shape and bound
// computations Lowering emits there) run once in the host launcher; an
// integer division or remainder by a loop invariant becomes a multiply-shift
// (Granlund-Montgomery) with the exact quotient of C's truncating `/` and
// `%`; nd_to_flat of a tuple known element by element is a closed form.
// Every fp64 operation is emitted as its own IEEE operation in source order
// (three-address form, CodegenC.hs:2293-2400) and the sources are compiled
// with -ffp-contract=off; integer + - * wrap (the oracle is built with -fwrapv).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <set>
#include <sstream>

#include "hydgen.hpp"

namespace hydgen {

namespace {

std::string sanitize(const std::string& v) {
    std::string out = "v_";
    for (char c : v) out += (isalnum((unsigned char)c) || c == '_') ? c : '_';
    return out;
}

std::string c_int(int64_t x) {
    if (x == INT64_MIN) return "((int64_t)(-9223372036854775807LL - 1))";
    return "((int64_t)" + std::to_string(x) + "LL)";
}

std::string c_double(double f) {
    if (f != f) return "__builtin_nan(\"\")";
    if (f == __builtin_inf()) return "__builtin_inf()";
    if (f == -__builtin_inf()) return "(-__builtin_inf())";
    char buf[64];
    snprintf(buf, sizeof buf, "%a", f);  // exact
    return std::string("(") + buf + ")";
}

const char* c_type(const CType& t) {
    switch (t.kind) {
        case CType::Int64:
        case CType::Bool: return "int64_t";
        case CType::Double: return "double";
        case CType::Tuple: return "hyd_tuple_t";
        default: throw Error("type " + t.show() + " cannot live in a kernel");
    }
}

bool is_float_op(BinOp op) {
    return op == BinOp::CAddF || op == BinOp::CSubF || op == BinOp::CMulF || op == BinOp::CDivF;
}
bool is_compare(BinOp op) {
    switch (op) {
        case BinOp::CEq: case BinOp::CNeq: case BinOp::CLt: case BinOp::CLe: case BinOp::CGt: case BinOp::CGe:
        case BinOp::CEqF: case BinOp::CNeqF: case BinOp::CLtF: case BinOp::CLeF: case BinOp::CGtF: case BinOp::CGeF:
        case BinOp::CAnd: case BinOp::COr: return true;
        default: return false;
    }
}
const char* c_binop(BinOp op) {
    switch (op) {
        case BinOp::CAdd: case BinOp::CAddF: return "+";
        case BinOp::CSub: case BinOp::CSubF: return "-";
        case BinOp::CMul: case BinOp::CMulF: return "*";
        case BinOp::CDiv: case BinOp::CDivF: return "/";
        case BinOp::CMod: return "%";
        case BinOp::CEq: case BinOp::CEqF: return "==";
        case BinOp::CNeq: case BinOp::CNeqF: return "!=";
        case BinOp::CLt: case BinOp::CLtF: return "<";
        case BinOp::CLe: case BinOp::CLeF: return "<=";
        case BinOp::CGt: case BinOp::CGtF: return ">";
        case BinOp::CGe: case BinOp::CGeF: return ">=";
        case BinOp::CAnd: return "&&";
        case BinOp::COr: return "||";
    }
    return "?";
}

// ---- the plan -----------------------------------------------------------------------
struct ArrayUse {
    CType elem;
    bool read = false, written = false, shape = false;
    bool loads_at_iter = true;   // every load indexes the kernel loop's iterator
    bool writes_at_iter = true;  // every write does
};

struct Triple {  // the scatter combine, Lowering.hs:3553-3557
    const Stmt* load = nullptr;  // old = dst[r]
    const Stmt* comb = nullptr;  // new = val + old
    const Stmt* write = nullptr; // dst[r] = new
    std::string dst;
    Atom route, val;
    bool in_loop = false;  // inside an inner loop of the body
};

struct Plan {
    enum Form { Map1D, Map2D, MapND, Reduce, ScatterInt, ScatterFloat } form = Map1D;
    const Stmt* kloop = nullptr;            // the distributed loop
    const Stmt* wrapper = nullptr;          // LoopReductionWrapper around it
    std::vector<std::string> iters;
    std::map<std::string, ArrayUse> arrays;
    std::map<std::string, CType> locals;    // assigned in the kernel body (and inner iterators)
    std::map<std::string, CType> scalars;   // scalar site arguments
    bool uses_exp = false;
    bool deferred = false;                  // compute-then-store rounds (1-D forms)
    // reduction
    std::string acc;
    CType acc_type;
    bool red_mul = false;
    // scatter
    Triple tri;
    // statements the launcher evaluates once (top level, one assignment,
    // operands that do not change over the loop), in order
    std::vector<const Stmt*> invariant_stmts;
    std::set<std::string> invariants;
    // tuples built once from components that do not change after (RTuple):
    // RNdToFlat / RProj of them are printed in closed form, as CodegenC.hs:
    // 2336-2346 prints RNdToFlat of a known tuple
    std::map<std::string, std::vector<Atom>> tuples;
    // integer divisors that do not change over the loop: / and % by them use
    // multiply-shift constants the launcher computes
    std::set<std::string> divisors;
};

const char* form_name(Plan::Form f) {
    switch (f) {
        case Plan::Map1D: return "map1d";
        case Plan::Map2D: return "map2d";
        case Plan::MapND: return "mapnd";
        case Plan::Reduce: return "reduce";
        case Plan::ScatterInt: return "scatter_i64";
        case Plan::ScatterFloat: return "scatter_f64";
    }
    return "?";
}

void for_each_stmt(const std::vector<Stmt>& ss, const std::function<void(const Stmt&)>& f) {
    for (const Stmt& s : ss) {
        f(s);
        if (s.kind == Stmt::SLoop || s.kind == Stmt::SIf || s.kind == Stmt::SParallelRegion) {
            for_each_stmt(s.body, f);
            for_each_stmt(s.els, f);
        }
    }
}

CType arg_type(const Site& site, const std::string& v, bool* found) {
    for (const auto& a : site.args)
        if (a.first == v) {
            *found = true;
            return a.second;
        }
    *found = false;
    return CType();
}

struct Analyzer {
    const Site& site;
    Plan& p;

    CType type_of(const std::string& v) {
        auto l = p.locals.find(v);
        if (l != p.locals.end()) return l->second;
        bool found;
        CType t = arg_type(site, v, &found);
        if (found) return t;
        auto t2 = site.types.find(v);
        if (t2 != site.types.end()) return t2->second;
        throw Error("site " + site.name + ": no type for variable " + v);
    }
    CType atom_type(const Atom& a) {
        switch (a.kind) {
            case Atom::Var: return type_of(a.var);
            case Atom::Int: case Atom::Unit: return CType::scalar(CType::Int64);
            case Atom::Float: return CType::scalar(CType::Double);
            case Atom::Bool: return CType::scalar(CType::Bool);
            default: throw Error("site " + site.name + ": string or vector atom in a kernel");
        }
    }
    CType rhs_type(const RHS& r) {
        switch (r.kind) {
            case RHS::RAtom: return atom_type(r.args[0]);
            case RHS::RBinOp:
                if (is_compare(r.bop)) return CType::scalar(CType::Bool);
                return CType::scalar(is_float_op(r.bop) ? CType::Double : CType::Int64);
            case RHS::RUnOp:
                switch (r.uop) {
                    case UnOp::CNot: return CType::scalar(CType::Bool);
                    case UnOp::CNeg: return atom_type(r.args[0]);
                    case UnOp::CIntOf: return CType::scalar(CType::Int64);
                    default: return CType::scalar(CType::Double);
                }
            case RHS::RArrayLoad: {
                CType t = atom_type(r.args[0]);
                if (!t.is_array() || !t.elem) throw Error("site " + site.name + ": load from a non-array");
                return *t.elem;
            }
            case RHS::RTuple: case RHS::RArrayShape: case RHS::RShapeInit: case RHS::RFlatToNd:
                return CType::scalar(CType::Tuple);
            case RHS::RProj: case RHS::RShapeSize: case RHS::RShapeLast: case RHS::RNdToFlat: case RHS::R2DToFlat:
                return CType::scalar(CType::Int64);
            default:
                throw Error("site " + site.name + ": right-hand side not supported in a kernel");
        }
    }

    void collect_locals(const std::vector<Stmt>& body) {
        for_each_stmt(body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign && !p.locals.count(s.var) && s.var != p.acc) {
                auto t = site.types.find(s.var);
                p.locals[s.var] = t != site.types.end() && t->second.kind != CType::Unknown ? t->second
                                                                                          : CType();
            }
            if (s.kind == Stmt::SLoop)
                for (const auto& it : s.loop.iters) p.locals[it] = CType::scalar(CType::Int64);
        });
        // types of locals the site's table lacks, in statement order (to a fixed point)
        for (int round = 0; round < 8; ++round) {
            bool missing = false;
            for_each_stmt(body, [&](const Stmt& s) {
                if (s.kind != Stmt::SAssign || s.var == p.acc) return;
                CType& t = p.locals[s.var];
                if (t.kind != CType::Unknown) return;
                try {
                    t = rhs_type(s.rhs);
                } catch (const Error&) {
                    missing = true;
                }
            });
            if (!missing) break;
        }
        for (const auto& l : p.locals)
            if (l.second.kind == CType::Unknown) throw Error("site " + site.name + ": cannot type " + l.first);
    }

    void note_arrays(const std::vector<Stmt>& body) {
        const std::string it = p.iters.size() == 1 ? p.iters[0] : std::string();
        auto arr = [&](const Atom& a) -> ArrayUse& {
            if (!a.is_var()) throw Error("site " + site.name + ": array operand is not a variable");
            bool found;
            CType t = arg_type(site, a.var, &found);
            if (!found || !t.is_array() || !t.elem)
                throw Error("site " + site.name + ": " + a.var + " is not an array argument of the site");
            ArrayUse& u = p.arrays[a.var];
            u.elem = *t.elem;
            return u;
        };
        for_each_stmt(body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign) {
                if (s.rhs.kind == RHS::RArrayLoad) {
                    ArrayUse& u = arr(s.rhs.args[0]);
                    u.read = true;
                    if (!(s.rhs.args[1].is_var() && s.rhs.args[1].var == it)) u.loads_at_iter = false;
                } else if (s.rhs.kind == RHS::RArrayShape) {
                    arr(s.rhs.args[0]).shape = true;
                } else {
                    for (const Atom& a : s.rhs.args)
                        if (a.is_var() && a.var != s.var) {
                            bool found;
                            CType t = arg_type(site, a.var, &found);
                            if (found && t.is_array())
                                throw Error("site " + site.name + ": array " + a.var + " used as a value");
                        }
                }
                if (s.rhs.kind == RHS::RUnOp && s.rhs.uop == UnOp::CExpF) p.uses_exp = true;
            } else if (s.kind == Stmt::SArrayWrite) {
                ArrayUse& u = arr(s.a0);
                u.written = true;
                if (!(s.a1.is_var() && s.a1.var == it)) u.writes_at_iter = false;
            } else if (s.kind == Stmt::SParallelRegion || s.kind == Stmt::SReturn) {
                throw Error("site " + site.name + ": parallel region or return inside an offloaded loop");
            }
        });
    }

    // The scatter combine anywhere in the body (not inside an inner loop).
    bool find_triple(const std::vector<Stmt>& ss, bool in_loop) {
        for (size_t k = 0; k < ss.size(); ++k) {
            const Stmt& a = ss[k];
            if (k + 2 < ss.size() && a.kind == Stmt::SAssign && a.rhs.kind == RHS::RArrayLoad &&
                a.rhs.args[0].is_var()) {
                const Stmt& b = ss[k + 1];
                const Stmt& c = ss[k + 2];
                const std::string dst = a.rhs.args[0].var;
                if (b.kind == Stmt::SAssign && b.rhs.kind == RHS::RBinOp &&
                    (b.rhs.bop == BinOp::CAdd || b.rhs.bop == BinOp::CAddF) && c.kind == Stmt::SArrayWrite &&
                    c.a0.is_var() && c.a0.var == dst && c.a2.is_var() && c.a2.var == b.var &&
                    c.a1.kind == a.rhs.args[1].kind && c.a1.var == a.rhs.args[1].var && c.a1.i == a.rhs.args[1].i) {
                    const Atom& x = b.rhs.args[0];
                    const Atom& y = b.rhs.args[1];
                    const bool xo = x.is_var() && x.var == a.var, yo = y.is_var() && y.var == a.var;
                    if (xo != yo) {
                        p.tri.load = &a;
                        p.tri.comb = &b;
                        p.tri.write = &c;
                        p.tri.dst = dst;
                        p.tri.route = a.rhs.args[1];
                        p.tri.val = xo ? y : x;
                        p.tri.in_loop = in_loop;
                        return true;
                    }
                }
            }
            if (a.kind == Stmt::SIf && (find_triple(a.body, in_loop) || find_triple(a.els, in_loop))) return true;
            if (a.kind == Stmt::SLoop && find_triple(a.body, true)) return true;
        }
        return false;
    }

    void run() {
        const Stmt& L = site.loop;
        if (L.loop.role == Role::ReductionWrapper && L.loop.iters.size() == 1 && L.loop.bounds[0].kind == IndexExpr::IConst &&
            L.loop.bounds[0].c == 1 && L.body.size() == 1 && L.body[0].kind == Stmt::SLoop && L.body[0].loop.has_red) {
            p.wrapper = &L;
            p.kloop = &L.body[0];
        } else {
            p.kloop = &L;
        }
        const LoopSpec& ks = p.kloop->loop;
        p.iters = ks.iters;
        if (ks.has_red) {
            if (ks.iters.size() != 1) throw Error("site " + site.name + ": multi-index reduction loop");
            p.form = Plan::Reduce;
            p.acc = ks.red_acc;
            p.red_mul = ks.red_mul;
            bool found;
            p.acc_type = arg_type(site, p.acc, &found);
            if (!found) throw Error("site " + site.name + ": the accumulator " + p.acc + " is not a site argument");
            if (!site.has_result || site.result.first != p.acc)
                throw Error("site " + site.name + ": a reduction site returns its accumulator");
        } else if (find_triple(p.kloop->body, false)) {
            if (ks.iters.size() != 1 || p.tri.in_loop)
                throw Error("site " + site.name + ": scatter combine in a multi-index or inner loop");
            bool found;
            CType dt = arg_type(site, p.tri.dst, &found);
            if (!found || !dt.is_array() || !dt.elem) throw Error("site " + site.name + ": scatter target");
            if (p.tri.comb->rhs.bop == BinOp::CAdd && dt.elem->kind == CType::Int64)
                p.form = Plan::ScatterInt;
            else if (p.tri.comb->rhs.bop == BinOp::CAddF && dt.elem->kind == CType::Double)
                p.form = Plan::ScatterFloat;
            else
                throw Error("site " + site.name + ": scatter combine does not match its target's type");
        } else {
            // a loop the parallelizer marked as a combining scatter without the
            // (+) / (+.) combine: its iterations race on the target
            if (ks.exec == LoopSpec::Parallel && ks.strategy != Strategy::Generic &&
                ks.strategy != Strategy::ScatterDirect)
                throw Error("site " + site.name + ": scatter with a combine other than + or +.");
            p.form = ks.iters.size() == 1 ? Plan::Map1D : ks.iters.size() == 2 ? Plan::Map2D : Plan::MapND;
        }
        for (const auto& a : site.args)
            if (!a.second.is_array() && a.first != p.acc) p.scalars[a.first] = a.second;
        collect_locals(p.kloop->body);
        if (p.wrapper)
            for (const auto& it : p.wrapper->loop.iters) p.locals[it] = CType::scalar(CType::Int64);
        for (const auto& it : p.iters) p.locals.erase(it);
        note_arrays(p.kloop->body);
        if (p.form == Plan::ScatterInt || p.form == Plan::ScatterFloat) {
            // the combine's own load and store of dst are not other uses
            int uses = 0;
            for_each_stmt(p.kloop->body, [&](const Stmt& s) {
                if (&s == p.tri.load || &s == p.tri.write) return;
                if (s.kind == Stmt::SAssign && !s.rhs.args.empty() && s.rhs.args[0].is_var() &&
                    s.rhs.args[0].var == p.tri.dst)
                    ++uses;
                if (s.kind == Stmt::SArrayWrite && s.a0.is_var() && s.a0.var == p.tri.dst) ++uses;
            });
            if (uses) throw Error("site " + site.name + ": the scatter target is used outside its combine");
        }
        // Deferred stores: top-level writes only, and no written array read.
        if (p.form == Plan::Map1D || p.form == Plan::MapND || p.form == Plan::Map2D) {
            bool ok = true;
            auto no_write = [&](const Stmt& t) {
                if (t.kind == Stmt::SArrayWrite) ok = false;
            };
            for (const Stmt& s : p.kloop->body)
                if (s.kind == Stmt::SIf || s.kind == Stmt::SLoop) {
                    for_each_stmt(s.body, no_write);
                    for_each_stmt(s.els, no_write);
                }
            for (const auto& a : p.arrays)
                if (a.second.written && a.second.read) ok = false;
            p.deferred = ok;
        }
        for (const auto& a : site.args)
            if (a.first != p.acc && p.locals.count(a.first))
                throw Error("site " + site.name + ": argument " + a.first + " is assigned inside the loop");
        check_free_vars();
        find_invariants();
    }

    // Invariant statements, known tuples and invariant divisors.
    void find_invariants() {
        std::map<std::string, int> count;
        for_each_stmt(p.kloop->body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign) count[s.var]++;
            if (s.kind == Stmt::SLoop)
                for (const auto& it : s.loop.iters) count[it] += 2;
        });
        std::set<std::string> iters(p.iters.begin(), p.iters.end());
        if (p.wrapper)
            for (const auto& it : p.wrapper->loop.iters) iters.insert(it);
        auto scalar_arg = [&](const std::string& v) {
            bool f;
            CType t = arg_type(site, v, &f);
            return f && !t.is_array() && v != p.acc;
        };
        auto inv_atom = [&](const Atom& a) {
            if (!a.is_var()) return a.kind == Atom::Int || a.kind == Atom::Float || a.kind == Atom::Bool ||
                                    a.kind == Atom::Unit;
            return p.invariants.count(a.var) > 0 || scalar_arg(a.var);
        };
        for (const Stmt& s : p.kloop->body) {
            if (s.kind != Stmt::SAssign || count[s.var] != 1 || s.var == p.acc || iters.count(s.var)) continue;
            bool ok = false;
            switch (s.rhs.kind) {
                case RHS::RAtom: case RHS::RBinOp: case RHS::RUnOp: case RHS::RProj: case RHS::RTuple:
                case RHS::RShapeSize: case RHS::RShapeInit: case RHS::RShapeLast: case RHS::R2DToFlat:
                case RHS::RNdToFlat: case RHS::RFlatToNd:
                    ok = std::all_of(s.rhs.args.begin(), s.rhs.args.end(), inv_atom);
                    break;
                case RHS::RArrayShape: ok = true; break;
                default: break;
            }
            if (ok) {
                p.invariants.insert(s.var);
                p.invariant_stmts.push_back(&s);
            }
        }
        // a tuple whose components cannot change after it is built
        auto fixed = [&](const Atom& a) {
            return !a.is_var() || p.invariants.count(a.var) || scalar_arg(a.var) || iters.count(a.var) ||
                   count[a.var] == 1;
        };
        for_each_stmt(p.kloop->body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign && s.rhs.kind == RHS::RTuple && count[s.var] == 1 &&
                std::all_of(s.rhs.args.begin(), s.rhs.args.end(), fixed))
                p.tuples[s.var] = s.rhs.args;
        });
        for_each_stmt(p.kloop->body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign && s.rhs.kind == RHS::RBinOp &&
                (s.rhs.bop == BinOp::CDiv || s.rhs.bop == BinOp::CMod) && s.rhs.args[1].is_var()) {
                const std::string& d = s.rhs.args[1].var;
                if (p.invariants.count(d) || scalar_arg(d)) {
                    CType t = type_of(d);
                    if (t.is_int_like()) p.divisors.insert(d);
                }
            }
        });
    }

    void check_free_vars() {
        std::set<std::string> bound(p.iters.begin(), p.iters.end());
        for (const auto& l : p.locals) bound.insert(l.first);
        if (!p.acc.empty()) bound.insert(p.acc);
        for (const auto& a : site.args) bound.insert(a.first);
        auto need = [&](const std::string& v) {
            if (!bound.count(v)) throw Error("site " + site.name + ": variable " + v + " is neither an argument nor local");
        };
        std::function<void(const IndexExpr&)> ie = [&](const IndexExpr& e) {
            if (e.kind == IndexExpr::IVar) need(e.var);
            for (const auto& k : e.kids) ie(k);
        };
        auto loop_exprs = [&](const LoopSpec& l) {
            for (const auto& b : l.bounds) ie(b);
            for (const auto& o : l.origins) ie(o);
            if (l.has_red) ie(l.red_init);
        };
        loop_exprs(site.loop.loop);
        if (p.wrapper) loop_exprs(p.kloop->loop);
        for_each_stmt(p.kloop->body, [&](const Stmt& s) {
            auto at = [&](const Atom& a) {
                if (a.is_var()) need(a.var);
            };
            if (s.kind == Stmt::SAssign)
                for (const Atom& a : s.rhs.args) at(a);
            if (s.kind == Stmt::SArrayWrite) at(s.a0), at(s.a1), at(s.a2);
            if (s.kind == Stmt::SIf) at(s.a0);
            if (s.kind == Stmt::SLoop) loop_exprs(s.loop);
        });
    }
};

Plan make_plan(const Site& site) {
    Plan p;
    Analyzer{site, p}.run();
    return p;
}

// A written array whose every write (outside inner loops) lands at
// (iter_0 + off_0, ..., iter_k + off_k) under one shape - a flat index of one
// iterator for the 1-D form - with loop-invariant offsets: the elements the
// site writes are then the box of its iteration space shifted by the offsets.
// Group launchers use the boxes to prove two sites' writes disjoint.
struct BoxWrite {
    std::vector<Atom> offs;  // per iteration dimension
    Atom shape;              // the flat index's shape (more than one dimension)
};

bool same_atom(const Atom& a, const Atom& b) {
    if (a.kind != b.kind) return false;
    if (a.kind == Atom::Var) return a.var == b.var;
    if (a.kind == Atom::Int) return a.i == b.i;
    return false;
}

std::map<std::string, BoxWrite> box_writes(const Site& site, const Plan& p) {
    std::map<std::string, BoxWrite> out;
    if (p.form != Plan::Map1D && p.form != Plan::Map2D && p.form != Plan::MapND) return out;
    const Stmt& loop = *p.kloop;
    std::map<std::string, int> count;
    for_each_stmt(loop.body, [&](const Stmt& st) {
        if (st.kind == Stmt::SAssign) count[st.var]++;
        if (st.kind == Stmt::SLoop)
            for (const auto& it : st.loop.iters) count[it] += 2;
    });
    std::map<std::string, const Stmt*> top;  // single top-level assignments
    for (const Stmt& st : loop.body)
        if (st.kind == Stmt::SAssign && count[st.var] == 1) top[st.var] = &st;
    auto inv = [&](const Atom& a) {
        if (a.kind == Atom::Int) return true;
        if (!a.is_var()) return false;
        if (p.invariants.count(a.var)) return true;
        auto it = p.scalars.find(a.var);
        return it != p.scalars.end() && a.var != p.acc;
    };
    // the offset of `a` from iterator d, when a = iter_d (+ invariant)
    auto iter_off = [&](const Atom& a, size_t d, Atom* off) {
        if (!a.is_var()) return false;
        if (a.var == p.iters[d]) {
            off->kind = Atom::Int;
            off->i = 0;
            return true;
        }
        auto it = top.find(a.var);
        if (it == top.end()) return false;
        const RHS& r = it->second->rhs;
        if (r.kind != RHS::RBinOp || r.bop != BinOp::CAdd) return false;
        if (r.args[0].is_var() && r.args[0].var == p.iters[d] && inv(r.args[1])) {
            *off = r.args[1];
            return true;
        }
        if (r.args[1].is_var() && r.args[1].var == p.iters[d] && inv(r.args[0])) {
            *off = r.args[0];
            return true;
        }
        return false;
    };
    const size_t nd = p.form == Plan::Map1D ? 1 : p.iters.size();
    std::set<std::string> bad;
    std::function<void(const std::vector<Stmt>&, int)> walk = [&](const std::vector<Stmt>& ss, int depth) {
        for (const Stmt& st : ss) {
            if (st.kind == Stmt::SArrayWrite) {
                const std::string& arr = st.a0.var;
                BoxWrite bw;
                bool ok = depth == 0 && st.a1.is_var();
                if (ok && nd == 1) {
                    Atom off;
                    ok = iter_off(st.a1, 0, &off);
                    bw.offs.push_back(off);
                } else if (ok) {
                    auto it = top.find(st.a1.var);
                    ok = it != top.end() && it->second->rhs.kind == RHS::RNdToFlat && it->second->rhs.args.size() == 2 &&
                         it->second->rhs.args[0].is_var() && inv(it->second->rhs.args[1]);
                    if (ok) {
                        bw.shape = it->second->rhs.args[1];
                        auto tt = top.find(it->second->rhs.args[0].var);
                        ok = tt != top.end() && tt->second->rhs.kind == RHS::RTuple && tt->second->rhs.args.size() == nd;
                        for (size_t d = 0; ok && d < nd; ++d) {
                            Atom off;
                            ok = iter_off(tt->second->rhs.args[d], d, &off);
                            bw.offs.push_back(off);
                        }
                    }
                }
                if (!ok) {
                    bad.insert(arr);
                } else {
                    auto prev = out.find(arr);
                    if (prev == out.end()) {
                        out[arr] = bw;
                    } else {
                        bool same = prev->second.offs.size() == bw.offs.size() && (nd == 1 || same_atom(prev->second.shape, bw.shape));
                        for (size_t d = 0; same && d < bw.offs.size(); ++d) same = same_atom(prev->second.offs[d], bw.offs[d]);
                        if (!same) bad.insert(arr);
                    }
                }
            } else if (st.kind == Stmt::SIf) {
                walk(st.body, depth);
                walk(st.els, depth);
            } else if (st.kind == Stmt::SLoop) {
                walk(st.body, depth + 1);
            }
        }
    };
    walk(loop.body, 0);
    for (const auto& b : bad) out.erase(b);
    (void)site;
    return out;
}

// A site whose body a group kernel can run: a map form without the exp
// table (its LDS prologue) in 256-thread workgroups.
bool groupable(const Plan& p, const Options& opt) {
    if (p.uses_exp) return false;
    if (p.form == Plan::Map2D) return true;
    return (p.form == Plan::Map1D || p.form == Plan::MapND) && opt.block == 256;
}

// ---- printing ----------------------------------------------------------------------
struct Printer {
    Printer(const Site& s, const Plan& plan, std::ostringstream& out) : site(s), p(plan), o(out) {
        for (const Stmt* st : p.invariant_stmts) hoisted.insert(st);
    }
    const Site& site;
    const Plan& p;
    std::ostringstream& o;
    std::set<const Stmt*> hoisted;
    bool host = false;          // printing for the launcher (host C++), not the kernel
    bool preloaded = false;     // streamed loads come from the round's 16-byte registers
    std::string acc_expr;       // how the kernel names the reduction accumulator
    enum WriteMode { Direct, Deferred } wmode = Direct;
    std::map<const Stmt*, int> write_slot;  // deferred writes: slot per top-level write
    mutable int loop_depth = 0;             // inner loops around the statement being printed
    enum ScatterMode { None, LdsBins, GlobalAtomics, Materialize } smode = None;

    std::string name(const std::string& v) const {
        if (!p.acc.empty() && v == p.acc && !acc_expr.empty()) return acc_expr;
        return sanitize(v);
    }
    std::string atom(const Atom& a) const {
        switch (a.kind) {
            case Atom::Var: return name(a.var);
            case Atom::Int: return c_int(a.i);
            case Atom::Float: return c_double(a.f);
            case Atom::Bool: return a.b ? "((int64_t)1)" : "((int64_t)0)";
            case Atom::Unit: return "((int64_t)0)";
            default: throw Error("site " + site.name + ": string or vector atom in a kernel");
        }
    }
    std::string index_expr(const IndexExpr& e) const {
        switch (e.kind) {
            case IndexExpr::IVar: return name(e.var);
            case IndexExpr::IConst: return c_int(e.c);
            case IndexExpr::IAdd: return "(" + index_expr(e.kids[0]) + " + " + index_expr(e.kids[1]) + ")";
            case IndexExpr::ISub: return "(" + index_expr(e.kids[0]) + " - " + index_expr(e.kids[1]) + ")";
            case IndexExpr::IMul: return "(" + index_expr(e.kids[0]) + " * " + index_expr(e.kids[1]) + ")";
            case IndexExpr::IDiv: return "(" + index_expr(e.kids[0]) + " / " + index_expr(e.kids[1]) + ")";
            case IndexExpr::IProj:
                if (e.kids[0].kind == IndexExpr::IVar)
                    return "(" + name(e.kids[0].var) + ".elems[" + std::to_string(e.c) + "])";
                [[fallthrough]];
            default: throw Error("site " + site.name + ": index expression not supported in a loop bound");
        }
    }
    const ArrayUse& use(const std::string& a) const { return p.arrays.at(a); }
    std::string arr(const std::string& a) const { return "hk_a_" + sanitize(a).substr(2); }
    bool one_d() const {
        return p.form == Plan::Map1D || p.form == Plan::Reduce || p.form == Plan::ScatterInt ||
               p.form == Plan::ScatterFloat;
    }
    bool stream_load(const std::string& a) const {
        const ArrayUse& u = use(a);
        return one_d() && u.read && u.loads_at_iter && !u.written;
    }
    bool stream_store(const std::string& a) const {
        const ArrayUse& u = use(a);
        return p.form == Plan::Map1D && u.written && u.writes_at_iter && !u.read;
    }
    bool known_tuple(const Atom& a) const { return a.is_var() && p.tuples.count(a.var); }

    std::string rhs(const RHS& r, const CType& target) const {
        const auto& A = r.args;
        switch (r.kind) {
            case RHS::RAtom: return atom(A[0]);
            case RHS::RBinOp: {
                const std::string a = atom(A[0]), b = atom(A[1]);
                if (is_float_op(r.bop)) return "(" + a + " " + c_binop(r.bop) + " " + b + ")";
                if (is_compare(r.bop)) return "((int64_t)(" + a + " " + c_binop(r.bop) + " " + b + "))";
                if (r.bop == BinOp::CAdd || r.bop == BinOp::CSub || r.bop == BinOp::CMul)
                    return "((int64_t)((uint64_t)" + a + " " + c_binop(r.bop) + " (uint64_t)" + b + "))";
                // CDiv, CMod: C truncation; by a divisor fixed over the loop, multiply-shift
                if (!host && A[1].is_var() && p.divisors.count(A[1].var))
                    return std::string(r.bop == BinOp::CDiv ? "hydk_cdiv(" : "hydk_cmod(") + a + ", hk_dc_" +
                           sanitize(A[1].var).substr(2) + ")";
                return "(" + a + " " + c_binop(r.bop) + " " + b + ")";
            }
            case RHS::RUnOp: {
                const std::string a = atom(A[0]);
                switch (r.uop) {
                    case UnOp::CNot: return "((int64_t)(!" + a + "))";
                    case UnOp::CNeg:
                        if (target.kind == CType::Double) return "(-" + a + ")";
                        return "((int64_t)(0ull - (uint64_t)" + a + "))";
                    case UnOp::CSqrt: return "__builtin_sqrt(" + a + ")";
                    case UnOp::CExpF: return host ? "hyd_exp(" + a + ")" : "hyd_exp_with(" + a + ", hk_exptab)";
                    case UnOp::CAbsF: return "__builtin_fabs(" + a + ")";
                    case UnOp::CFloorF: return "__builtin_floor(" + a + ")";
                    case UnOp::CCeilF: return "__builtin_ceil(" + a + ")";
                    case UnOp::CFloatOf: return "((double)" + a + ")";
                    case UnOp::CIntOf: return "((int64_t)" + a + ")";
                    default:
                        throw Error("site " + site.name +
                                    ": log/sin/cos/erf have no glibc-exact device form; the loop stays on the host");
                }
            }
            case RHS::RArrayLoad: {
                const std::string& a = A[0].var;
                const std::string ix = atom(A[1]);
                if (stream_load(a)) {
                    if (preloaded) {
                        const std::string reg = "hk_p_" + sanitize(a).substr(2) + "[hk_u >> 1]";
                        return "((hk_u & 1) ? " + reg + ".y : " + reg + ".x)";
                    }
                    return "__builtin_nontemporal_load(&" + arr(a) + "[" + ix + "])";
                }
                return arr(a) + "[" + ix + "]";
            }
            case RHS::RTuple: {
                if (A.size() > 8) throw Error("site " + site.name + ": tuple of more than 8 components");
                std::string s = "hydk_tuple(" + std::to_string(A.size());
                for (const Atom& a : A) s += ", " + atom(a);
                return s + ")";
            }
            case RHS::RProj:
                if (known_tuple(A[0])) {
                    const auto& comp = p.tuples.at(A[0].var);
                    if (r.proj >= 0 && r.proj < (int64_t)comp.size()) return atom(comp[(size_t)r.proj]);
                }
                return "(" + atom(A[0]) + ".elems[" + std::to_string(r.proj) + "])";
            case RHS::RArrayShape:
                return host ? sanitize(A[0].var) + "->shape" : "hk_shape_" + sanitize(A[0].var).substr(2);
            case RHS::RShapeSize: return "hydk_shape_size(" + atom(A[0]) + ")";
            case RHS::RShapeInit: return "hydk_shape_init(" + atom(A[0]) + ")";
            case RHS::RShapeLast: return "hydk_shape_last(" + atom(A[0]) + ")";
            case RHS::RFlatToNd: return "hydk_flat_to_nd(" + atom(A[0]) + ", " + atom(A[1]) + ")";
            case RHS::RNdToFlat: {
                if (A[0].kind == Atom::Unit) return "((int64_t)0)";
                if (known_tuple(A[0]) && p.tuples.at(A[0].var).size() >= 2) {
                    // row-major: ((e0 * s1 + e1) * s2 + e2) ..., wrapping
                    const auto& comp = p.tuples.at(A[0].var);
                    const std::string shp = atom(A[1]);
                    std::string acc = "(uint64_t)" + atom(comp[0]);
                    for (size_t k = 1; k < comp.size(); ++k)
                        acc = "(" + acc + " * (uint64_t)" + shp + ".elems[" + std::to_string(k) + "] + (uint64_t)" +
                              atom(comp[k]) + ")";
                    return "((int64_t)" + acc + ")";
                }
                return "hydk_nd_to_flat(" + atom(A[0]) + ", " + atom(A[1]) + ")";
            }
            case RHS::R2DToFlat: return "((int64_t)((uint64_t)" + atom(A[0]) + " * (uint64_t)" + atom(A[1]) + "))";
            default: throw Error("site " + site.name + ": right-hand side not supported in a kernel");
        }
    }

    CType var_type(const std::string& v) const {
        if (!p.acc.empty() && v == p.acc) return p.acc_type;
        auto l = p.locals.find(v);
        if (l != p.locals.end()) return l->second;
        throw Error("site " + site.name + ": assignment to non-local " + v);
    }

    void line(int ind, const std::string& s) const { o << std::string((size_t)ind * 4, ' ') << s << "\n"; }

    void declare_locals(int ind) const {
        for (const auto& l : p.locals)
            if (!p.invariants.count(l.first)) line(ind, std::string(c_type(l.second)) + " " + name(l.first) + ";");
    }

    // The launcher's evaluation of the invariant statements (host C++).
    void hoisted_stmts(int ind) {
        host = true;
        for (const Stmt* s : p.invariant_stmts) {
            const CType t = var_type(s->var);
            line(ind, std::string("const ") + c_type(t) + " " + name(s->var) + " = (" + c_type(t) + ")" +
                          rhs(s->rhs, t) + ";");
        }
        host = false;
    }

    void stmts(const std::vector<Stmt>& ss, int ind, std::set<std::string>& assigned) const {
        for (size_t k = 0; k < ss.size(); ++k) {
            const Stmt& s = ss[k];
            if (hoisted.count(&s)) continue;  // a kernel parameter
            if (smode != None && &s == p.tri.load) {
                scatter_contribution(ind);
                k += 2;  // the combine and its write
                continue;
            }
            switch (s.kind) {
                case Stmt::SAssign: {
                    const CType t = var_type(s.var);
                    line(ind, name(s.var) + " = (" + c_type(t) + ")" + rhs(s.rhs, t) + ";");
                    assigned.insert(s.var);
                    break;
                }
                case Stmt::SArrayWrite: {
                    const std::string& a = s.a0.var;
                    const std::string et = c_type(use(a).elem);
                    if (wmode == Deferred) {
                        const int slot = write_slot.at(&s);
                        line(ind, "hk_wi" + std::to_string(slot) + "[hk_u] = " + atom(s.a1) + ";");
                        line(ind, "hk_wv" + std::to_string(slot) + "[hk_u] = (" + et + ")" + atom(s.a2) + ";");
                    } else if (stream_store(a)) {
                        line(ind, "__builtin_nontemporal_store((" + et + ")" + atom(s.a2) + ", &" + arr(a) + "[" +
                                      atom(s.a1) + "]);");
                    } else {
                        line(ind, arr(a) + "[" + atom(s.a1) + "] = (" + et + ")" + atom(s.a2) + ";");
                    }
                    break;
                }
                case Stmt::SIf: {
                    line(ind, "if (" + atom(s.a0) + ") {");
                    std::set<std::string> a1 = assigned, a2 = assigned;
                    stmts(s.body, ind + 1, a1);
                    if (!s.els.empty()) {
                        line(ind, "} else {");
                        stmts(s.els, ind + 1, a2);
                    }
                    line(ind, "}");
                    for (const auto& v : a1)
                        if (a2.count(v)) assigned.insert(v);
                    break;
                }
                case Stmt::SLoop: {
                    const LoopSpec& l = s.loop;
                    // CodegenC.hs:1305-1330: the accumulator starts at rsInit
                    // unless it was assigned before the loop
                    if (l.has_red && !assigned.count(l.red_acc))
                        line(ind, name(l.red_acc) + " = (" + c_type(var_type(l.red_acc)) + ")" + index_expr(l.red_init) + ";");
                    for (size_t d = 0; d < l.iters.size(); ++d) {
                        const std::string it = name(l.iters[d]);
                        const std::string lo = index_expr(l.origins[d]);
                        line(ind + (int)d, "for (" + it + " = " + lo + "; " + it + " < " + lo + " + " +
                                               index_expr(l.bounds[d]) + "; ++" + it + ") {");
                    }
                    std::set<std::string> inner = assigned;
                    ++loop_depth;
                    stmts(s.body, ind + (int)l.iters.size(), inner);
                    --loop_depth;
                    for (size_t d = l.iters.size(); d-- > 0;) line(ind + (int)d, "}");
                    if (l.has_red) assigned.insert(l.red_acc);
                    break;
                }
                case Stmt::SBreak:
                    // a break of the offloaded loop itself has no parallel meaning
                    if (loop_depth == 0) throw Error("site " + site.name + ": break out of the offloaded loop");
                    line(ind, "break;");
                    break;
                default: throw Error("site " + site.name + ": statement not supported in a kernel");
            }
        }
    }

    void scatter_contribution(int ind) const {
        const std::string r = atom(p.tri.route), v = atom(p.tri.val);
        switch (smode) {
            case LdsBins:
                line(ind, "{ const int64_t hk_r = " + r + ";");
                line(ind, "  if ((uint64_t)hk_r < (uint64_t)hk_ndst) hydk_lds_add(&hk_mine[hk_r], (uint64_t)" + v + "); }");
                break;
            case GlobalAtomics:
                line(ind, "{ const int64_t hk_r = " + r + ";");
                line(ind, "  if ((uint64_t)hk_r < (uint64_t)hk_ndst) hydk_global_add((unsigned long long*)&" +
                              arr(p.tri.dst) + "[hk_r], (uint64_t)" + v + "); }");
                break;
            case Materialize:
                line(ind, "hk_sidx[hk_e] = " + r + "; hk_sval[hk_e] = (double)" + v + "; hk_skeep[hk_e] = 1;");
                break;
            default: break;
        }
    }
};

// ---- kernel and launcher text -----------------------------------------------------------
struct Param {
    std::string ctype, name, launcher_arg;
};

std::vector<Param> kernel_params(const Site& site, const Plan& p) {
    std::vector<Param> ps;
    for (const auto& a : site.args) {
        const std::string& v = a.first;
        if (a.second.is_array()) {
            auto it = p.arrays.find(v);
            if (it == p.arrays.end()) continue;  // not touched by the kernel loop
            const std::string base = sanitize(v).substr(2);
            ps.push_back({std::string(c_type(it->second.elem)) + "*", "hk_a_" + base, "hk_a_" + base});
            if (it->second.shape) ps.push_back({"hyd_tuple_t", "hk_shape_" + base, sanitize(v) + "->shape"});
        } else if (v != p.acc) {
            ps.push_back({c_type(a.second), sanitize(v), sanitize(v)});
        }
    }
    for (const Stmt* s : p.invariant_stmts) {
        auto it = p.locals.find(s->var);
        ps.push_back({c_type(it->second), sanitize(s->var), sanitize(s->var)});
    }
    for (const std::string& d : p.divisors) {
        const std::string base = sanitize(d).substr(2);
        ps.push_back({"hydk_divc_t", "hk_dc_" + base, "hk_dc_" + base});
    }
    return ps;
}

std::string param_list(const std::vector<Param>& ps) {
    std::string s;
    for (const auto& q : ps) s += (s.empty() ? "" : ", ") + q.ctype + " " + q.name;
    return s;
}
std::string arg_list(const std::vector<Param>& ps) {
    std::string s;
    for (const auto& q : ps) s += (s.empty() ? "" : ", ") + q.launcher_arg;
    return s;
}
std::string name_list(const std::vector<Param>& ps, const std::string& prefix = "") {
    std::string s;
    for (const auto& q : ps) s += (s.empty() ? "" : ", ") + prefix + q.name;
    return s;
}
// The index-range arguments of a map form's kernel, in order.
std::vector<std::string> map_idx_vars(const Plan& p) {
    std::vector<std::string> v;
    if (p.form == Plan::Map2D) return {"hk_lo0", "hk_n0", "hk_lo1", "hk_n1"};
    if (p.form == Plan::MapND) {
        for (size_t d = 0; d < p.iters.size(); ++d) {
            v.push_back("hk_lo" + std::to_string(d));
            v.push_back("hk_n" + std::to_string(d));
        }
        v.push_back("hk_n");
        return v;
    }
    return {"hk_lo", "hk_n"};
}
std::string map_idx_names(const Plan& p, const std::string& prefix = "") {
    std::string s;
    for (const auto& v : map_idx_vars(p)) s += (s.empty() ? "" : ", ") + prefix + v;
    return s;
}

void exp_table_prologue(std::ostringstream& o, const Plan& p) {
    if (!p.uses_exp) return;
    o << "    __shared__ uint64_t hk_exptab[256];\n"
         "    for (int hk_t = threadIdx.x; hk_t < 256; hk_t += blockDim.x) hk_exptab[hk_t] = hyd_exp_tab_dev[hk_t];\n"
         "    __syncthreads();\n";
}

// One element of a 1-D or N-D form: declarations, the wrapper's index, the
// kernel index, the body.
void element_body(Printer& pr, const Plan& p, int ind, const std::string& index_expr) {
    pr.declare_locals(ind);
    if (p.wrapper)
        for (size_t d = 0; d < p.wrapper->loop.iters.size(); ++d)
            pr.line(ind, pr.name(p.wrapper->loop.iters[d]) + " = " + pr.index_expr(p.wrapper->loop.origins[d]) + ";");
    if (p.form == Plan::MapND) {
        // flat element -> indices, last fastest
        pr.line(ind, "int64_t hk_rest = " + index_expr + ";");
        for (size_t d = p.iters.size(); d-- > 0;) {
            const std::string n = "hk_n" + std::to_string(d);
            pr.line(ind, "const int64_t " + sanitize(p.iters[d]) + " = hk_lo" + std::to_string(d) + " + hk_rest % " + n + ";");
            pr.line(ind, "hk_rest /= " + n + ";");
        }
    } else {
        pr.line(ind, "const int64_t " + sanitize(p.iters[0]) + " = hk_lo + " + index_expr + ";");
    }
    if (p.form == Plan::ScatterFloat) pr.line(ind, "hk_skeep[hk_e] = 0;");
    std::set<std::string> assigned;
    pr.stmts(p.kloop->body, ind, assigned);
}

const char* vec2_type(const CType& t) { return t.kind == CType::Double ? "hydk_v2d" : "hydk_v2i"; }

// The streaming element loop of the 1-D and N-D forms.  A round gives each
// thread `U` elements; with `paired` (U even, a 1-D form) elements 2q and
// 2q+1 of the thread are adjacent, so the streamed arrays move in 16-byte
// accesses when PAIRS (a template parameter the launcher sets when they are
// 16-byte aligned) and in two 8-byte accesses otherwise, with the same
// element-to-thread assignment either way (a reduction's order does not
// depend on alignment).
void stream_loop(std::ostringstream& o, Printer& pr, const Plan& p, int U, int B, bool paired, bool deferred,
                 const std::string& bx = "blockIdx.x", const std::string& gx = "gridDim.x") {
    const int T = U * B;
    const std::string emap = paired ? "hk_base + 2 * ((int64_t)(hk_u >> 1) * " + std::to_string(B) +
                                          " + threadIdx.x) + (hk_u & 1)"
                                    : "hk_base + (int64_t)hk_u * " + std::to_string(B) + " + threadIdx.x";
    std::vector<std::string> sload, sstore;  // streamed arrays
    for (const auto& a : p.arrays) {
        if (pr.stream_load(a.first)) sload.push_back(a.first);
        if (deferred && pr.stream_store(a.first)) sstore.push_back(a.first);
    }
    auto full_tile = [&](bool pairs_io) {
        if (pairs_io)
            for (const auto& a : sload) {
                const std::string reg = "hk_p_" + sanitize(a).substr(2);
                const char* vt = vec2_type(pr.use(a).elem);
                o << "            " << vt << " " << reg << "[" << U / 2 << "];\n";
                o << "#pragma unroll\n";
                o << "            for (int hk_q = 0; hk_q < " << U / 2 << "; ++hk_q)\n";
                o << "                " << reg << "[hk_q] = __builtin_nontemporal_load((const " << vt << "*)&" << pr.arr(a)
                  << "[hk_lo + hk_base + 2 * ((int64_t)hk_q * " << B << " + threadIdx.x)]);\n";
            }
        if (deferred) {
            int slot = 0;
            pr.write_slot.clear();
            for (const Stmt& s : p.kloop->body)
                if (s.kind == Stmt::SArrayWrite) {
                    pr.write_slot[&s] = slot;
                    o << "            int64_t hk_wi" << slot << "[" << U << "];\n";
                    o << "            " << c_type(p.arrays.at(s.a0.var).elem) << " hk_wv" << slot << "[" << U << "];\n";
                    ++slot;
                }
            pr.wmode = Printer::Deferred;
        }
        pr.preloaded = pairs_io;
        o << "#pragma unroll\n";
        o << "            for (int hk_u = 0; hk_u < " << U << "; ++hk_u) {\n";
        o << "                const int64_t hk_e = " << emap << ";\n";
        element_body(pr, p, 4, "hk_e");
        o << "            }\n";
        pr.preloaded = false;
        if (deferred) {
            pr.wmode = Printer::Direct;
            for (const auto& ws : pr.write_slot) {
                const std::string& a = ws.first->a0.var;
                const std::string sl = std::to_string(ws.second);
                const bool streamed = std::find(sstore.begin(), sstore.end(), a) != sstore.end();
                if (streamed && pairs_io) {
                    const char* vt = vec2_type(pr.use(a).elem);
                    o << "#pragma unroll\n";
                    o << "            for (int hk_q = 0; hk_q < " << U / 2 << "; ++hk_q) {\n";
                    o << "                " << vt << " hk_w2 = {hk_wv" << sl << "[2 * hk_q], hk_wv" << sl << "[2 * hk_q + 1]};\n";
                    o << "                __builtin_nontemporal_store(hk_w2, (" << vt << "*)&" << pr.arr(a) << "[hk_wi" << sl
                      << "[2 * hk_q]]);\n";
                    o << "            }\n";
                } else {
                    o << "#pragma unroll\n";
                    o << "            for (int hk_u = 0; hk_u < " << U << "; ++hk_u)\n";
                    if (streamed)
                        o << "                __builtin_nontemporal_store(hk_wv" << sl << "[hk_u], &" << pr.arr(a) << "[hk_wi"
                          << sl << "[hk_u]]);\n";
                    else
                        o << "                " << pr.arr(a) << "[hk_wi" << sl << "[hk_u]] = hk_wv" << sl << "[hk_u];\n";
                }
            }
        }
    };
    o << "    for (int64_t hk_base = (int64_t)" << bx << " * " << T << "; hk_base < hk_n; hk_base += (int64_t)" << gx << " * "
      << T << ") {\n";
    o << "        if (hk_base + " << T << " <= hk_n) {\n";
    if (paired && (!sload.empty() || !sstore.empty())) {
        o << "          if constexpr (PAIRS) {\n";
        full_tile(true);
        o << "          } else {\n";
        full_tile(false);
        o << "          }\n";
    } else {
        full_tile(false);
    }
    o << "        } else {\n";
    o << "            for (int hk_u = 0; hk_u < " << U << "; ++hk_u) {\n";
    o << "                const int64_t hk_e = " << emap << ";\n";
    o << "                if (hk_e >= hk_n) continue;\n";
    element_body(pr, p, 4, "hk_e");
    o << "            }\n";
    o << "        }\n";
    o << "    }\n";
}

// Map forms: the kernel's arguments are computed by hk_prep_<site> into a
// struct (hk_m_<site>, passed to the kernel) plus host-side facts
// (hk_x_<site>: whether there is work, the 16-byte form, the write boxes), so
// that the site's launcher and a group launcher share them.
void emit_map_launcher(std::ostringstream& o, const Site& site, const Plan& p, const Options& opt, Printer& pr,
                       const std::vector<Param>& ps, bool has_pairs, const std::vector<std::string>& streamed,
                       const std::string& lparams, const std::string& largs) {
    const std::string M = "hk_m_" + site.name, X = "hk_x_" + site.name, K = "hk_kernel_" + site.name;
    const std::vector<std::string> iv = map_idx_vars(p);
    const std::map<std::string, BoxWrite> boxes = box_writes(site, p);
    const size_t nd = p.form == Plan::Map1D ? 1 : p.iters.size();
    const int U = std::max(1, opt.unroll), B = opt.block;
    o << "struct " << M << " {\n";
    for (const auto& q : ps) o << "    " << q.ctype << " " << q.name << ";\n";
    for (const auto& v : iv) o << "    int64_t " << v << ";\n";
    o << "    int64_t hk_gx, hk_gy;\n};\n";
    o << "struct " << X << " {\n    int hk_live, hk_pairs;\n";
    for (const auto& b : boxes) {
        const std::string base = sanitize(b.first).substr(2);
        o << "    int64_t hk_bl_" << base << "[" << nd << "], hk_bh_" << base << "[" << nd << "];\n";
        o << "    hyd_tuple_t hk_bs_" << base << ";\n";
    }
    o << "};\n";
    o << "static inline " << M << " hk_prep_" << site.name << "(" << lparams << (lparams.empty() ? "" : ", ") << X
      << "* hk_x) {\n";
    o << "    const char* hk_site = \"" << site.name << "\";\n";
    o << "    (void)hk_site;\n";
    o << "    " << M << " hk_m;\n";
    o << "    memset(&hk_m, 0, sizeof hk_m);\n";
    o << "    memset(hk_x, 0, sizeof *hk_x);\n";
    for (const auto& a : p.arrays) {
        const std::string base = sanitize(a.first).substr(2);
        o << "    " << c_type(a.second.elem) << "* hk_a_" << base << " = (" << c_type(a.second.elem)
          << "*)hydk_device_data(" << sanitize(a.first) << ");\n";
    }
    const LoopSpec& ks = p.kloop->loop;
    std::string nonempty;
    if (p.form == Plan::Map2D) {
        o << "    const int64_t hk_lo0 = " << pr.index_expr(ks.origins[0]) << ", hk_n0 = " << pr.index_expr(ks.bounds[0]) << ";\n";
        o << "    const int64_t hk_lo1 = " << pr.index_expr(ks.origins[1]) << ", hk_n1 = " << pr.index_expr(ks.bounds[1]) << ";\n";
        nonempty = "hk_n0 > 0 && hk_n1 > 0";
    } else if (p.form == Plan::MapND) {
        o << "    int64_t hk_n = 1;\n";
        for (size_t d = 0; d < p.iters.size(); ++d) {
            o << "    const int64_t hk_lo" << d << " = " << pr.index_expr(ks.origins[d]) << ", hk_n" << d << " = "
              << pr.index_expr(ks.bounds[d]) << ";\n";
            o << "    hk_n = hk_n" << d << " > 0 ? hk_n * hk_n" << d << " : 0;\n";
        }
        nonempty = "hk_n > 0";
    } else {
        o << "    const int64_t hk_lo = " << pr.index_expr(ks.origins[0]) << ", hk_n = " << pr.index_expr(ks.bounds[0]) << ";\n";
        nonempty = "hk_n > 0";
    }
    o << "    if (" << nonempty << ") {\n";
    pr.hoisted_stmts(2);
    for (const std::string& d : p.divisors)
        o << "        const hydk_divc_t hk_dc_" << sanitize(d).substr(2) << " = hydk_make_divc(" << sanitize(d) << ");\n";
    if (has_pairs) {
        o << "        const bool hk_pairs = true";
        for (const auto& a : streamed) o << " && hydk_aligned16(hk_a_" << sanitize(a).substr(2) << " + hk_lo)";
        o << ";\n";
    }
    if (p.form == Plan::Map2D) {
        const int R = p.deferred ? opt.rows : 1;
        const int64_t rows_per_band = (int64_t)R * opt.band_rounds;
        o << "        const int64_t hk_gx = (hk_n1 + 255) / 256;\n";
        o << "        const int64_t hk_gy = std::min<int64_t>(65535, (hk_n0 + " << rows_per_band - 1 << ") / " << rows_per_band
          << ");\n";
        o << "        if (hk_gx > 0x7fffffff) hydk_fail(hk_site, \"inner extent too large\");\n";
    } else {
        // flat: the grid sized by the data, one round per workgroup (per_cu 0)
        o << "        const int64_t hk_gx = hydk_stream_grid(hk_n, " << U * B << ", "
          << (opt.flat_maps ? 0 : opt.waves_per_cu * 64 / B) << "), hk_gy = 1;\n";
    }
    for (const auto& q : ps) o << "        hk_m." << q.name << " = " << q.launcher_arg << ";\n";
    for (const auto& v : iv) o << "        hk_m." << v << " = " << v << ";\n";
    o << "        hk_m.hk_gx = hk_gx;\n        hk_m.hk_gy = hk_gy;\n";
    o << "        hk_x->hk_live = 1;\n        hk_x->hk_pairs = " << (has_pairs ? "hk_pairs" : "0") << ";\n";
    for (const auto& b : boxes) {
        const std::string base = sanitize(b.first).substr(2);
        for (size_t d = 0; d < nd; ++d) {
            const std::string lo = nd == 1 ? "hk_lo" : "hk_lo" + std::to_string(d);
            const std::string n = nd == 1 ? "hk_n" : "hk_n" + std::to_string(d);
            o << "        hk_x->hk_bl_" << base << "[" << d << "] = " << lo << " + (int64_t)" << pr.atom(b.second.offs[d]) << ";\n";
            o << "        hk_x->hk_bh_" << base << "[" << d << "] = hk_x->hk_bl_" << base << "[" << d << "] + " << n << ";\n";
        }
        if (nd > 1) o << "        hk_x->hk_bs_" << base << " = " << pr.atom(b.second.shape) << ";\n";
    }
    o << "    }\n";
    o << "    return hk_m;\n}\n\n";

    o << "extern \"C\" int64_t " << site.name << "(" << lparams << ") {\n";
    o << "    const char* hk_site = \"" << site.name << "\";\n";
    o << "    hipStream_t hk_s = hydk_stream();\n";
    o << "    " << X << " hk_x;\n";
    o << "    const " << M << " hk_m = hk_prep_" << site.name << "(" << largs << (largs.empty() ? "" : ", ") << "&hk_x);\n";
    o << "    if (hk_x.hk_live) {\n";
    const std::string grid = p.form == Plan::Map2D ? "dim3((unsigned)hk_m.hk_gx, (unsigned)hk_m.hk_gy)" : "(unsigned)hk_m.hk_gx";
    const std::string block = p.form == Plan::Map2D ? "256" : std::to_string(B);
    const std::string pl = name_list(ps, "hk_m.");
    const std::string args = pl + (pl.empty() ? "" : ", ") + map_idx_names(p, "hk_m.");
    if (has_pairs) {
        o << "        if (hk_x.hk_pairs) " << K << "<true><<<" << grid << ", " << block << ", 0, hk_s>>>(" << args << ");\n";
        o << "        else " << K << "<false><<<" << grid << ", " << block << ", 0, hk_s>>>(" << args << ");\n";
    } else {
        o << "        " << K << "<<<" << grid << ", " << block << ", 0, hk_s>>>(" << args << ");\n";
    }
    o << "        hydk_check_launch(hk_site);\n";
    o << "    }\n";
    for (const auto& a : p.arrays)
        if (a.second.written) o << "    hydk_device_written(" << sanitize(a.first) << ");\n";
    o << "    return 0;\n}\n\n";
}

void emit_site(std::ostringstream& o, const Site& site, const Options& opt) {
    const Plan p = make_plan(site);
    const std::string K = "hk_kernel_" + site.name;
    const int U = std::max(1, opt.unroll);
    const int B = opt.block;
    const std::vector<Param> ps = kernel_params(site, p);
    const std::string plist = param_list(ps);
    const std::string sep = plist.empty() ? "" : ", ";
    Printer pr(site, p, o);
    const bool one_d = pr.one_d();
    const bool paired = one_d && U % 2 == 0;
    std::vector<std::string> streamed;  // arrays the PAIRS variant moves in 16-byte accesses
    for (const auto& a : p.arrays)
        if (pr.stream_load(a.first) || (p.deferred && pr.stream_store(a.first))) streamed.push_back(a.first);
    const bool has_pairs = paired && !streamed.empty();
    const std::string tmpl = has_pairs ? "template <bool PAIRS>\n" : "";

    o << "// ---- " << site.name << ": " << describe_plan(site, opt) << "\n";
    std::string idx_params = p.form == Plan::MapND ? "" : "int64_t hk_lo, int64_t hk_n";
    if (p.form == Plan::MapND) {
        for (size_t d = 0; d < p.iters.size(); ++d)
            idx_params += std::string(d ? ", " : "") + "int64_t hk_lo" + std::to_string(d) + ", int64_t hk_n" +
                          std::to_string(d);
        idx_params += ", int64_t hk_n";
    }

    // Map forms: the body is a device function of the workgroup's grid
    // position, so that a group launcher can run several sites' bodies in one
    // launch; the site's own kernel calls it with blockIdx / gridDim.
    const std::string Kb = "hk_body_" + site.name;
    const std::string gparams = "unsigned hk_bx, unsigned hk_by, unsigned hk_gx, unsigned hk_gy";
    auto map_wrapper = [&](int block, const std::string& iparams, const std::string& inames) {
        o << tmpl << "__global__ __launch_bounds__(" << block << ") void " << K << "(" << plist << sep << iparams << ") {\n";
        o << "    " << Kb << (has_pairs ? "<PAIRS>" : "") << "(" << name_list(ps) << sep << inames
          << ", blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);\n";
        o << "}\n\n";
    };
    if (p.form == Plan::Map1D || p.form == Plan::MapND) {
        o << tmpl << "__device__ __forceinline__ void " << Kb << "(" << plist << sep << idx_params << ", " << gparams
          << ") {\n";
        exp_table_prologue(o, p);
        stream_loop(o, pr, p, U, B, paired, p.deferred, "hk_bx", "hk_gx");
        o << "}\n\n";
        map_wrapper(B, idx_params, map_idx_names(p));
    } else if (p.form == Plan::Map2D) {
        // 256-thread workgroups spanning 256 consecutive columns (the inner,
        // contiguous index: each row access of a workgroup is 2 KiB of one
        // row); workgroup row y owns the band of rows [y*band, (y+1)*band),
        // band a multiple of R, and walks it R rows per round; with deferred
        // stores a round computes its R rows before storing them, so loads of
        // one address in a round are loaded once.  The launcher makes bands
        // `band_rounds` rounds long (default one): on MI355X many short
        // workgroups, balanced by the dispatcher, ran the 8192 x 8192 heat
        // step in 244 us against 292 us for one resident round of long bands
        const int R = p.deferred ? opt.rows : 1;
        o << "__device__ __forceinline__ void " << Kb << "(" << plist << sep
          << "int64_t hk_lo0, int64_t hk_n0, int64_t hk_lo1, int64_t hk_n1, " << gparams << ") {\n";
        exp_table_prologue(o, p);
        o << "    const int64_t hk_c = (int64_t)hk_bx * 256 + threadIdx.x;\n";
        o << "    if (hk_c >= hk_n1) return;\n";
        o << "    const int64_t " << sanitize(p.iters[1]) << " = hk_lo1 + hk_c;\n";
        o << "    const int64_t hk_band = ((hk_n0 + " << R << " - 1) / " << R << " + (int64_t)hk_gy - 1) / (int64_t)hk_gy * "
          << R << ";\n";
        o << "    const int64_t hk_r1 = min(hk_n0, ((int64_t)hk_by + 1) * hk_band);\n";
        o << "    for (int64_t hk_r = (int64_t)hk_by * hk_band; hk_r < hk_r1; hk_r += " << R << ") {\n";
        auto row_body = [&](int ind) {
            pr.declare_locals(ind);
            o << std::string((size_t)ind * 4, ' ') << "const int64_t " << sanitize(p.iters[0]) << " = hk_lo0 + hk_r + hk_u;\n";
            std::set<std::string> assigned;
            pr.stmts(p.kloop->body, ind, assigned);
        };
        if (R > 1) {
            o << "        if (hk_r + " << R << " <= hk_r1) {\n";
            int slot = 0;
            pr.write_slot.clear();
            for (const Stmt& st : p.kloop->body)
                if (st.kind == Stmt::SArrayWrite) {
                    pr.write_slot[&st] = slot;
                    o << "            int64_t hk_wi" << slot << "[" << R << "];\n";
                    o << "            " << c_type(p.arrays.at(st.a0.var).elem) << " hk_wv" << slot << "[" << R << "];\n";
                    ++slot;
                }
            pr.wmode = Printer::Deferred;
            o << "#pragma unroll\n";
            o << "            for (int hk_u = 0; hk_u < " << R << "; ++hk_u) {\n";
            row_body(4);
            o << "            }\n";
            pr.wmode = Printer::Direct;
            for (const auto& ws : pr.write_slot) {
                const std::string sl = std::to_string(ws.second);
                o << "#pragma unroll\n";
                o << "            for (int hk_u = 0; hk_u < " << R << "; ++hk_u) " << pr.arr(ws.first->a0.var) << "[hk_wi" << sl
                  << "[hk_u]] = hk_wv" << sl << "[hk_u];\n";
            }
            o << "        } else {\n";
            o << "            for (int hk_u = 0; hk_u < " << R << " && hk_r + hk_u < hk_r1; ++hk_u) {\n";
            row_body(4);
            o << "            }\n";
            o << "        }\n";
        } else {
            o << "        {\n";
            o << "            const int hk_u = 0;\n";
            row_body(3);
            o << "        }\n";
        }
        o << "    }\n}\n\n";
        map_wrapper(256, "int64_t hk_lo0, int64_t hk_n0, int64_t hk_lo1, int64_t hk_n1", "hk_lo0, hk_n0, hk_lo1, hk_n1");
    } else if (p.form == Plan::Reduce) {
        const std::string T = c_type(p.acc_type);
        const bool fl = p.acc_type.kind == CType::Double;
        const std::string ident = p.red_mul ? (fl ? "1.0" : "((int64_t)1)") : (fl ? "(-0.0)" : "((int64_t)0)");
        const std::string comb = p.red_mul ? (fl ? "hydk_mul_f64" : "hydk_mul_i64") : (fl ? "hydk_add_f64" : "hydk_add_i64");
        o << tmpl << "__global__ __launch_bounds__(" << B << ") void " << K << "(" << plist << sep << idx_params << ", " << T
          << "* hk_partials) {\n";
        exp_table_prologue(o, p);
        o << "    " << T << " hk_acc[" << U << "];\n";
        o << "#pragma unroll\n    for (int hk_u = 0; hk_u < " << U << "; ++hk_u) hk_acc[hk_u] = " << ident << ";\n";
        pr.acc_expr = "hk_acc[hk_u]";
        stream_loop(o, pr, p, U, B, paired, false);
        pr.acc_expr.clear();
        // fixed combine: the thread's accumulators in order, the wave by xor
        // butterfly, the waves in order
        o << "    " << T << " hk_v = hk_acc[0];\n";
        o << "#pragma unroll\n    for (int hk_u = 1; hk_u < " << U << "; ++hk_u) hk_v = " << comb << "(hk_v, hk_acc[hk_u]);\n";
        o << "    hk_v = hydk_wave_reduce(hk_v, " << comb << "_op());\n";
        o << "    __shared__ " << T << " hk_waves[" << B / 64 << "];\n";
        o << "    if ((threadIdx.x & 63) == 0) hk_waves[threadIdx.x >> 6] = hk_v;\n";
        o << "    __syncthreads();\n";
        o << "    if (threadIdx.x == 0) {\n";
        o << "        " << T << " hk_w = hk_waves[0];\n";
        o << "        for (int hk_k = 1; hk_k < " << B / 64 << "; ++hk_k) hk_w = " << comb << "(hk_w, hk_waves[hk_k]);\n";
        o << "        hk_partials[blockIdx.x] = hk_w;\n";
        o << "    }\n}\n\n";
        // the final fold over the workgroup partials
        o << "__global__ __launch_bounds__(256) void " << K << "_final(const " << T << "* hk_partials, int64_t hk_np, " << T
          << " hk_init, " << T << "* hk_out) {\n";
        o << "    " << T << " hk_v = " << ident << ";\n";
        o << "    for (int64_t hk_k = threadIdx.x; hk_k < hk_np; hk_k += 256) hk_v = " << comb << "(hk_v, hk_partials[hk_k]);\n";
        o << "    hk_v = hydk_wave_reduce(hk_v, " << comb << "_op());\n";
        o << "    __shared__ " << T << " hk_waves[4];\n";
        o << "    if ((threadIdx.x & 63) == 0) hk_waves[threadIdx.x >> 6] = hk_v;\n";
        o << "    __syncthreads();\n";
        o << "    if (threadIdx.x == 0) {\n";
        o << "        " << T << " hk_w = " << comb << "(" << comb << "(hk_waves[0], hk_waves[1]), " << comb
          << "(hk_waves[2], hk_waves[3]));\n";
        o << "        hk_out[0] = " << comb << "(hk_init, hk_w);\n";
        o << "    }\n}\n\n";
    } else if (p.form == Plan::ScatterInt) {
        // LDS-private bins, merged through spread copies of a zeroed scratch
        o << tmpl << "__global__ __launch_bounds__(" << B << ") void " << K << "_lds(" << plist << sep << idx_params
          << ", int64_t hk_ndst, int hk_copies, int hk_stride, int hk_ncopy, unsigned long long* hk_spread) {\n";
        exp_table_prologue(o, p);
        o << "    extern __shared__ unsigned long long hk_bins[];\n";
        o << "    for (int hk_t = threadIdx.x; hk_t < hk_copies * hk_stride; hk_t += " << B << ") hk_bins[hk_t] = 0;\n";
        o << "    __syncthreads();\n";
        o << "    unsigned long long* hk_mine = hk_bins + (threadIdx.x % hk_copies) * hk_stride;\n";
        pr.smode = Printer::LdsBins;
        stream_loop(o, pr, p, U, B, paired, false);
        o << "    __syncthreads();\n";
        o << "    unsigned long long* hk_dst_copy = hk_spread + (size_t)(blockIdx.x % hk_ncopy) * hk_ndst;\n";
        o << "    for (int64_t hk_b = threadIdx.x; hk_b < hk_ndst; hk_b += " << B << ") {\n";
        o << "        unsigned long long hk_s = 0;\n";
        o << "        for (int hk_c = 0; hk_c < hk_copies; ++hk_c) hk_s += hk_bins[hk_c * hk_stride + hk_b];\n";
        o << "        if (hk_s) hydk_global_add(&hk_dst_copy[hk_b], hk_s);\n";
        o << "    }\n}\n\n";
        o << tmpl << "__global__ __launch_bounds__(" << B << ") void " << K << "_global(" << plist << sep << idx_params
          << ", int64_t hk_ndst) {\n";
        exp_table_prologue(o, p);
        pr.smode = Printer::GlobalAtomics;
        stream_loop(o, pr, p, U, B, paired, false);
        o << "}\n\n";
        pr.smode = Printer::None;
    } else if (p.form == Plan::ScatterFloat) {
        o << tmpl << "__global__ __launch_bounds__(" << B << ") void " << K << "(" << plist << sep << idx_params
          << ", int64_t* hk_sidx, double* hk_sval, int64_t* hk_skeep) {\n";
        exp_table_prologue(o, p);
        pr.smode = Printer::Materialize;
        stream_loop(o, pr, p, U, B, paired, false);
        pr.smode = Printer::None;
        o << "}\n\n";
    }

    // ---- launcher
    const std::string ret = site.has_result ? c_type(site.result.second) : "int64_t";
    std::string lparams, largs;
    for (const auto& a : site.args) {
        lparams += lparams.empty() ? "" : ", ";
        lparams += a.second.is_array() ? "hyd_array_t* " + sanitize(a.first)
                                       : std::string(c_type(a.second)) + " " + sanitize(a.first);
        largs += (largs.empty() ? "" : ", ") + sanitize(a.first);
    }
    if (p.form == Plan::Map1D || p.form == Plan::Map2D || p.form == Plan::MapND) {
        emit_map_launcher(o, site, p, opt, pr, ps, has_pairs, streamed, lparams, largs);
        return;
    }
    o << "extern \"C\" " << ret << " " << site.name << "(" << lparams << ") {\n";
    o << "    const char* hk_site = \"" << site.name << "\";\n";
    o << "    hipStream_t hk_s = hydk_stream();\n";
    for (const auto& a : p.arrays) {
        const std::string base = sanitize(a.first).substr(2);
        o << "    " << c_type(a.second.elem) << "* hk_a_" << base << " = (" << c_type(a.second.elem)
          << "*)hydk_device_data(" << sanitize(a.first) << ");\n";
    }
    if (p.form == Plan::Reduce) o << "    " << c_type(p.acc_type) << " hk_result = " << sanitize(p.acc) << ";\n";
    // reductions and scatters (map forms: emit_map_launcher): one index
    const LoopSpec& ks = p.kloop->loop;
    std::string nonempty;
    o << "    const int64_t hk_lo = " << pr.index_expr(ks.origins[0]) << ", hk_n = " << pr.index_expr(ks.bounds[0]) << ";\n";
    nonempty = "hk_n > 0";
    if (p.form == Plan::ScatterInt || p.form == Plan::ScatterFloat) {
        o << "    const int64_t hk_ndst = hyd_shape_size(" << sanitize(p.tri.dst) << "->shape);\n";
        nonempty += " && hk_ndst > 0";
    }
    o << "    if (" << nonempty << ") {\n";
    // the loop's invariant statements, once, and the divisors' constants
    pr.hoisted_stmts(2);
    for (const std::string& d : p.divisors)
        o << "        const hydk_divc_t hk_dc_" << sanitize(d).substr(2) << " = hydk_make_divc(" << sanitize(d) << ");\n";
    std::string targs;  // template argument of the paired kernels
    if (has_pairs) {
        o << "        const bool hk_pairs = true";
        for (const auto& a : streamed) o << " && hydk_aligned16(hk_a_" << sanitize(a).substr(2) << " + hk_lo)";
        o << ";\n";
    }
    auto launch = [&](const std::string& kern, const std::string& grid, const std::string& block, const std::string& lds,
                      const std::string& args, int ind) {
        const std::string pad((size_t)ind * 4, ' ');
        if (has_pairs) {
            o << pad << "if (hk_pairs) " << kern << "<true><<<" << grid << ", " << block << ", " << lds << ", hk_s>>>(" << args
              << ");\n";
            o << pad << "else " << kern << "<false><<<" << grid << ", " << block << ", " << lds << ", hk_s>>>(" << args << ");\n";
        } else {
            o << pad << kern << "<<<" << grid << ", " << block << ", " << lds << ", hk_s>>>(" << args << ");\n";
        }
        o << pad << "hydk_check_launch(hk_site);\n";
    };
    const std::string al = arg_list(ps);
    {
        const std::string A = al + sep + "hk_lo, hk_n";
        const bool scatter = p.form == Plan::ScatterInt || p.form == Plan::ScatterFloat;
        o << "        const int64_t hk_grid = hydk_stream_grid(hk_n, " << U * B << ", "
          << (scatter ? 2 : 1) * opt.waves_per_cu * 64 / B << ");\n";
        if (p.form == Plan::Reduce) {
            const std::string T = c_type(p.acc_type);
            o << "        " << T << "* hk_part = (" << T << "*)hydk_scratch(0, (hk_grid + 1) * 8, hk_s);\n";
            launch(K, "(unsigned)hk_grid", std::to_string(B), "0", A + ", hk_part", 2);
            o << "        " << K << "_final<<<1, 256, 0, hk_s>>>(hk_part, hk_grid, " << sanitize(p.acc) << ", hk_part + hk_grid);\n";
            o << "        hydk_check_launch(hk_site);\n";
            o << "        hydk_read_back(&hk_result, hk_part + hk_grid, sizeof hk_result, hk_s, hk_site);\n";
        } else if (p.form == Plan::ScatterInt) {
            o << "        if (hk_ndst <= 8192) {\n";
            o << "            const int hk_stride = (int)(hk_ndst | 1);\n";
            o << "            const int hk_copies = std::max(1, std::min(16, 8192 / hk_stride));\n";
            o << "            const int hk_ncopy = (int)std::max<int64_t>(1, std::min<int64_t>(32, 4096 / hk_ndst));\n";
            o << "            unsigned long long* hk_spread = (unsigned long long*)hydk_zero_scratch(hk_s);\n";
            launch(K + "_lds", "(unsigned)hk_grid", std::to_string(B), "(size_t)hk_copies * hk_stride * 8",
                   A + ", hk_ndst, hk_copies, hk_stride, hk_ncopy, hk_spread", 3);
            o << "            const int hk_used = (int)std::min<int64_t>(hk_ncopy, hk_grid);\n";
            o << "            int hk_lanes = 1;\n";
            o << "            while (hk_lanes < hk_used) hk_lanes <<= 1;\n";
            o << "            hydk_spread_fold<<<(unsigned)((hk_ndst * hk_lanes + 255) / 256), 256, 0, hk_s>>>(hk_spread, hk_used, "
                 "hk_lanes, hk_ndst, (unsigned long long*)hk_a_"
              << sanitize(p.tri.dst).substr(2) << ");\n";
            o << "            hydk_check_launch(hk_site);\n";
            o << "        } else {\n";
            launch(K + "_global", "(unsigned)hk_grid", std::to_string(B), "0", A + ", hk_ndst", 3);
            o << "        }\n";
        } else if (p.form == Plan::ScatterFloat) {
            o << "        unsigned char* hk_m = (unsigned char*)hydk_scratch(2, hk_n * 24, hk_s);\n";
            o << "        int64_t* hk_sidx = (int64_t*)hk_m;\n";
            o << "        double* hk_sval = (double*)(hk_m + hk_n * 8);\n";
            o << "        int64_t* hk_skeep = (int64_t*)(hk_m + hk_n * 16);\n";
            launch(K, "(unsigned)hk_grid", std::to_string(B), "0", A + ", hk_sidx, hk_sval, hk_skeep", 2);
            o << "        hydk_check(hydhip_scatter_guarded_add_f64(hk_sidx, hk_sval, hk_skeep, hk_n, hk_a_"
              << sanitize(p.tri.dst).substr(2) << ", hk_ndst, hk_s), hk_site);\n";
        }
    }
    o << "    }\n";
    for (const auto& a : p.arrays)
        if (a.second.written) o << "    hydk_device_written(" << sanitize(a.first) << ");\n";
    o << "    return " << (p.form == Plan::Reduce ? "hk_result" : "0") << ";\n";
    o << "}\n\n";
}

}  // namespace

std::string describe_plan(const Site& site, const Options& opt) {
    const Plan p = make_plan(site);
    std::ostringstream o;
    o << "form=" << form_name(p.form) << " iters=" << p.iters.size();
    const bool one_d = p.form == Plan::Map1D || p.form == Plan::Reduce || p.form == Plan::ScatterInt ||
                       p.form == Plan::ScatterFloat;
    if (p.form != Plan::Map2D) o << " block=" << opt.block << " unroll=" << opt.unroll;
    else
        o << " block=256x1 rows=" << (p.deferred ? opt.rows : 1) << " grid.y="
          << "rows/" << (p.deferred ? opt.rows : 1) * opt.band_rounds;
    if (one_d && opt.unroll % 2 == 0) o << " paired=1";
    if (p.wrapper) o << " wrapper=1";
    if (p.form == Plan::Map1D || p.form == Plan::MapND) o << " deferred=" << (p.deferred ? 1 : 0);
    if (p.uses_exp) o << " exp_lds=1";
    if (!p.invariants.empty()) {
        o << " hoisted=";
        bool f = true;
        for (const Stmt* s : p.invariant_stmts) o << (f ? "" : ",") << s->var, f = false;
    }
    if (!p.divisors.empty()) {
        o << " divisors=";
        bool f = true;
        for (const auto& d : p.divisors) o << (f ? "" : ",") << d, f = false;
    }
    if (!p.tuples.empty()) o << " tuples=" << p.tuples.size();
    o << " arrays=";
    bool first = true;
    for (const auto& a : p.arrays) {
        o << (first ? "" : ",") << a.first << ":" << (a.second.read ? "r" : "") << (a.second.written ? "w" : "");
        const bool sl = one_d && a.second.loads_at_iter && !a.second.written && a.second.read;
        const bool ss = p.form == Plan::Map1D && a.second.writes_at_iter && !a.second.read && a.second.written;
        if (sl || ss) o << "(stream)";
        first = false;
    }
    if (p.form == Plan::Reduce) o << " op=" << (p.red_mul ? "mul" : "add") << " acc=" << p.acc;
    if (p.form == Plan::ScatterInt || p.form == Plan::ScatterFloat) o << " dst=" << p.tri.dst;
    return o.str();
}

namespace {

// The 16-byte (PAIRS) variant exists for a 1-D map whose streamed arrays
// are all 16-byte aligned at run time; a group kernel runs the plain one.
bool site_has_pairs(const Plan& p, const Options& opt) {
    if (p.form != Plan::Map1D || std::max(1, opt.unroll) % 2 != 0) return false;
    for (const auto& a : p.arrays) {
        const ArrayUse& u = a.second;
        if (u.read && u.loads_at_iter && !u.written) return true;
        if (p.deferred && u.written && u.writes_at_iter && !u.read) return true;
    }
    return false;
}

void emit_group(std::ostringstream& o, const Group& g, const std::map<std::string, const Site*>& by_name,
                const Options& opt) {
    std::vector<const Site*> ms;
    for (const auto& m : g.members) ms.push_back(by_name.at(m));
    std::vector<Plan> plans;
    bool fusable = true;
    for (const Site* m : ms) {
        plans.push_back(make_plan(*m));
        fusable = fusable && groupable(plans.back(), opt);
    }
    std::string lparams;
    {
        // the union of the members' arguments (group_site checked their types)
        std::map<std::string, CType> seen;
        std::vector<std::pair<std::string, CType>> args;
        for (const Site* m : ms)
            for (const auto& a : m->args)
                if (seen.emplace(a.first, a.second).second) args.push_back(a);
        for (const auto& a : args) {
            lparams += lparams.empty() ? "" : ", ";
            lparams += a.second.is_array() ? "hyd_array_t* " + sanitize(a.first)
                                           : std::string(c_type(a.second)) + " " + sanitize(a.first);
        }
    }
    auto member_args = [&](const Site* m) {
        std::string s;
        for (const auto& a : m->args) s += (s.empty() ? "" : ", ") + sanitize(a.first);
        return s;
    };
    const size_t N = ms.size();
    o << "// ---- group " << g.name << ":";
    for (const auto& m : g.members) o << " " << m;
    o << (fusable ? " (one launch when independent)" : " (in order: a member is not a 256-thread map form)") << "\n";
    if (fusable) {
        o << "__global__ __launch_bounds__(256) void hk_group_" << g.name << "(";
        for (size_t k = 0; k < N; ++k) o << "hk_m_" << ms[k]->name << " hk_m" << k << ", ";
        for (size_t k = 0; k < N; ++k) o << "int64_t hk_e" << k << (k + 1 < N ? ", " : "");
        o << ") {\n";
        o << "    const int64_t hk_b = blockIdx.x;\n";
        for (size_t k = 0; k < N; ++k) {
            const std::string mk = "hk_m" + std::to_string(k);
            const std::vector<Param> ps = kernel_params(*ms[k], plans[k]);
            const std::string pl = name_list(ps, mk + ".");
            o << "    if (hk_b < hk_e" << k << ") {\n";
            o << "        const unsigned hk_l = (unsigned)(hk_b" << (k ? " - hk_e" + std::to_string(k - 1) : std::string()) << ");\n";
            o << "        hk_body_" << ms[k]->name << (site_has_pairs(plans[k], opt) ? "<false>" : "") << "(" << pl
              << (pl.empty() ? "" : ", ") << map_idx_names(plans[k], mk + ".") << ", hk_l % (unsigned)" << mk
              << ".hk_gx, hk_l / (unsigned)" << mk << ".hk_gx, (unsigned)" << mk << ".hk_gx, (unsigned)" << mk << ".hk_gy);\n";
            o << "        return;\n    }\n";
        }
        o << "}\n\n";
    }
    o << "extern \"C\" int64_t " << g.name << "(" << lparams << ") {\n";
    if (!fusable) {
        for (const Site* m : ms) o << "    (void)" << m->name << "(" << member_args(m) << ");\n";
        o << "    return 0;\n}\n\n";
        return;
    }
    o << "    const char* hk_site = \"" << g.name << "\";\n";
    o << "    hipStream_t hk_s = hydk_stream();\n";
    size_t naccs = 0;
    for (size_t k = 0; k < N; ++k) {
        const std::string ma = member_args(ms[k]);
        o << "    hk_x_" << ms[k]->name << " hk_x" << k << ";\n";
        o << "    const hk_m_" << ms[k]->name << " hk_m" << k << " = hk_prep_" << ms[k]->name << "(" << ma
          << (ma.empty() ? "" : ", ") << "&hk_x" << k << ");\n";
        naccs += plans[k].arrays.size();
    }
    o << "    hydk_acc_t hk_acc[" << std::max<size_t>(naccs, 1) << "];\n";
    o << "    int hk_na = 0;\n";
    for (size_t k = 0; k < N; ++k) {
        const Plan& p = plans[k];
        const auto boxes = box_writes(*ms[k], p);
        const size_t nd = p.form == Plan::Map1D ? 1 : p.iters.size();
        const std::string xk = "hk_x" + std::to_string(k), mk = "hk_m" + std::to_string(k);
        o << "    if (" << xk << ".hk_live) {\n";
        for (const auto& a : p.arrays) {
            if (!a.second.read && !a.second.written) continue;
            const std::string base = sanitize(a.first).substr(2);
            const bool boxed = a.second.written && boxes.count(a.first);
            o << "        hk_acc[hk_na++] = hydk_acc(" << k << ", " << mk << ".hk_a_" << base << ", hyd_shape_size("
              << sanitize(a.first) << "->shape) * (int64_t)sizeof(" << c_type(a.second.elem) << "), "
              << (a.second.read ? 1 : 0) << ", " << (a.second.written ? 1 : 0) << ", " << (boxed ? 1 : 0) << ", " << nd
              << ", " << (boxed ? xk + ".hk_bl_" + base : std::string("nullptr")) << ", "
              << (boxed ? xk + ".hk_bh_" + base : std::string("nullptr")) << ", "
              << (boxed ? "&" + xk + ".hk_bs_" + base : std::string("nullptr")) << ");\n";
        }
        o << "    }\n";
    }
    for (size_t k = 0; k < N; ++k) {
        const std::string xk = "hk_x" + std::to_string(k), mk = "hk_m" + std::to_string(k);
        o << "    const int64_t hk_e" << k << " = " << (k ? "hk_e" + std::to_string(k - 1) + " + " : std::string()) << "("
          << xk << ".hk_live ? " << mk << ".hk_gx * " << mk << ".hk_gy : 0);\n";
    }
    o << "    const bool hk_fuse = hydk_accesses_independent(hk_acc, hk_na) && hk_e" << N - 1 << " <= 0x7fffffff && (";
    for (size_t k = 0; k < N; ++k) o << (k ? " + " : "") << "sizeof(hk_m" << k << ")";
    o << ") <= 3072;\n";
    o << "    if (!hk_fuse) {\n";
    for (const Site* m : ms) o << "        (void)" << m->name << "(" << member_args(m) << ");\n";
    o << "        return 0;\n    }\n";
    o << "    if (hk_e" << N - 1 << " > 0) {\n";
    o << "        hk_group_" << g.name << "<<<(unsigned)hk_e" << N - 1 << ", 256, 0, hk_s>>>(";
    for (size_t k = 0; k < N; ++k) o << "hk_m" << k << ", ";
    for (size_t k = 0; k < N; ++k) o << "hk_e" << k << (k + 1 < N ? ", " : "");
    o << ");\n";
    o << "        hydk_check_launch(hk_site);\n    }\n";
    std::set<std::string> written;
    for (size_t k = 0; k < N; ++k)
        for (const auto& a : plans[k].arrays)
            if (a.second.written && written.insert(a.first).second)
                o << "    hydk_device_written(" << sanitize(a.first) << ");\n";
    o << "    return 0;\n}\n\n";
}

}  // namespace

Site group_site(const Group& g, const std::vector<Site>& sites) {
    Site gs;
    gs.name = g.name;
    if (sites.empty()) return gs;
    std::map<std::string, const Site*> by_name;
    for (const Site& s : sites) by_name[s.name] = &s;
    std::map<std::string, CType> seen;
    for (const auto& m : g.members) {
        auto it = by_name.find(m);
        if (it == by_name.end()) throw Error("group " + g.name + ": no site named " + m);
        for (const auto& a : it->second->args) {
            auto s = seen.find(a.first);
            if (s == seen.end()) {
                seen[a.first] = a.second;
                gs.args.push_back(a);
            } else if (s->second.show() != a.second.show()) {
                throw Error("group " + g.name + ": argument " + a.first + " has two types");
            }
        }
    }
    return gs;
}

std::string emit_sites(const std::vector<Site>& sites, const std::vector<Group>& groups, const Options& opt,
                       const std::string& origin) {
    std::ostringstream o;
    o << "// Generated by hydgen from " << origin << ": one kernel and launcher per offloaded loop.\n"
         "// Compile with hipcc --offload-arch=gfx950 -O3 -ffp-contract=off and link with -lhydhip.\n"
         "#include \"hydgen_rt.h\"\n\n";
    std::set<std::string> names;
    std::map<std::string, const Site*> by_name;
    for (const Site& s : sites) {
        if (!names.insert(s.name).second) throw Error("two sites named " + s.name);
        by_name[s.name] = &s;
        emit_site(o, s, opt);
    }
    std::set<std::string> grouped;
    for (const Group& g : groups) {
        if (!names.insert(g.name).second) throw Error("group " + g.name + ": the name is taken");
        if (g.members.empty()) throw Error("group " + g.name + ": no members");
        for (const auto& m : g.members) {
            if (!by_name.count(m)) throw Error("group " + g.name + ": no site named " + m);
            if (!grouped.insert(m).second) throw Error("group " + g.name + ": site " + m + " is in two groups");
        }
        (void)group_site(g, sites);  // the members' arguments agree
        emit_group(o, g, by_name, opt);
    }
    return o.str();
}

std::string emit_signatures(const std::vector<Site>& sites) {
    auto tag = [](const CType& t) -> std::string {
        if (t.is_array()) return "array";
        if (t.kind == CType::Tuple) return "tuple";
        if (t.kind == CType::Double) return "f64";
        if (t.kind == CType::Int64 || t.kind == CType::Bool) return "i64";
        throw Error("type " + t.show() + " has no launcher form");
    };
    std::ostringstream o;
    for (const Site& s : sites) {
        o << s.name << " " << (s.has_result ? tag(s.result.second) : std::string("i64")) << " :";
        for (const auto& a : s.args) o << " " << tag(a.second) << ":" << a.first;
        o << "\n";
    }
    return o.str();
}

std::string emit_header(const std::vector<Site>& sites) {
    std::ostringstream o;
    o << "/* Generated by hydgen: the launchers of the offloaded loops. */\n"
         "#ifndef HYDGEN_SITES_H\n#define HYDGEN_SITES_H\n#include <stdint.h>\n#include \"hydrangea_runtime.h\"\n"
         "#ifdef __cplusplus\nextern \"C\" {\n#endif\n";
    for (const Site& s : sites) {
        const std::string ret = s.has_result ? c_type(s.result.second) : "int64_t";
        std::string ps;
        for (const auto& a : s.args) {
            ps += ps.empty() ? "" : ", ";
            ps += a.second.is_array() ? "hyd_array_t* " + sanitize(a.first)
                                      : std::string(c_type(a.second)) + " " + sanitize(a.first);
        }
        o << ret << " " << s.name << "(" << (ps.empty() ? "void" : ps) << ");\n";
    }
    o << "#ifdef __cplusplus\n}\n#endif\n#endif\n";
    return o.str();
}

}  // namespace hydgen

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
#include <memory>
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
// the register array of column offset dc in the band form
std::string dc_name(int64_t dc) { return dc < 0 ? "m" + std::to_string(-dc) : dc > 0 ? "p" + std::to_string(dc) : "0"; }

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
    // The temporal stencil form: loads of `tile_array` read the workgroup's
    // LDS tile hk_src (the flat index variables in `tile_pos` print as tile
    // offsets of their (row, column)), and the write of `out_array` sets
    // hk_val, the cell's new value.
    std::string tile_array, out_array;
    const std::map<std::string, std::pair<Atom, Atom>>* tile_pos = nullptr;
    // ... or, in the band form, from registers: the load at (row + dr,
    // col + dc) of the cell in row hk_k of the wave's band is
    // hk_x<dc>[hk_k + reg_rr + dr]
    const std::map<std::string, std::pair<int64_t, int64_t>>* reg_pos = nullptr;
    int reg_rr = 0;

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
                if (!tile_array.empty() && a == tile_array) {
                    if (reg_pos && A[1].is_var() && reg_pos->count(A[1].var)) {
                        const auto& d = reg_pos->at(A[1].var);
                        return "hk_x" + dc_name(d.second) + "[hk_k + " + std::to_string(reg_rr + d.first) + "]";
                    }
                    return "hk_src[" + ix + "]";
                }
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
                    if (reg_pos && reg_pos->count(s.var)) {
                        // the load reads a register: no index
                    } else if (tile_pos && tile_pos->count(s.var)) {
                        const auto& rc = tile_pos->at(s.var);
                        line(ind, name(s.var) + " = (int64_t)((int)(" + atom(rc.first) + " - hk_R0) * hk_LW + (int)(" +
                                      atom(rc.second) + " - hk_C0));");
                    } else {
                        line(ind, name(s.var) + " = (" + c_type(t) + ")" + rhs(s.rhs, t) + ";");
                    }
                    assigned.insert(s.var);
                    break;
                }
                case Stmt::SArrayWrite: {
                    const std::string& a = s.a0.var;
                    const std::string et = c_type(use(a).elem);
                    if (!out_array.empty() && a == out_array) {
                        line(ind, "hk_val = (" + et + ")" + atom(s.a2) + ";");
                    } else if (wmode == Deferred) {
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

// ---- an iterate of a stencil step (a LoopIterate site) -----------------------------------
//
// The site's loop is the whole `iterate`: Lowering.hs:840-857 lowers it to a
// serial LoopIterate loop whose body is the step (its statements, one array
// allocated for the new state, the step's loops writing it) followed by
// cur = next.  The launcher evaluates the step's scalar statements once (they
// depend on the state's shape, never on its elements or the counter), finds
// the step's loops the branches select, and runs the steps:
//   * temporal form: when the selected loops are one run of 2-index maps that
//     write the new state at (i + a, j + b), read the old one only at
//     (i + c, j + d) or at fixed rows/columns near the cell, and exactly tile
//     the grid, one launch advances K steps over an LDS tile with a K-cell
//     halo per unit of radius (each cell runs the body of the loop whose box
//     holds it, reading its neighbours from the tile);
//   * otherwise every step calls the loops' own launchers in order, into a
//     zeroed array (the reference allocates a fresh, calloc'd one per step).
// Two HBM buffers ping-pong; the result is the last one written (the input
// itself after zero steps, as the reference's loop leaves it).
namespace {

void rhs_vars(const RHS& r, std::set<std::string>& out) {
    for (const Atom& a : r.args)
        if (a.is_var()) out.insert(a.var);
}
void index_vars(const IndexExpr& e, std::set<std::string>& out) {
    if (e.kind == IndexExpr::IVar) out.insert(e.var);
    for (const auto& k : e.kids) index_vars(k, out);
}
// Variables statements read and assign (loop iterators count as assigned).
void stmt_vars(const std::vector<Stmt>& ss, std::set<std::string>& used, std::set<std::string>& assigned) {
    for_each_stmt(ss, [&](const Stmt& s) {
        switch (s.kind) {
            case Stmt::SAssign: rhs_vars(s.rhs, used); assigned.insert(s.var); break;
            case Stmt::SArrayWrite:
                for (const Atom* a : {&s.a0, &s.a1, &s.a2})
                    if (a->is_var()) used.insert(a->var);
                break;
            case Stmt::SIf: case Stmt::SReturn:
                if (s.a0.is_var()) used.insert(s.a0.var);
                break;
            case Stmt::SLoop:
                for (const auto& b : s.loop.bounds) index_vars(b, used);
                for (const auto& b : s.loop.origins) index_vars(b, used);
                if (s.loop.has_red) index_vars(s.loop.red_init, used);
                for (const auto& it : s.loop.iters) assigned.insert(it);
                break;
            default: break;
        }
    });
}

// a*i + b*j + sum(c_v * v) + k over the kernel iterators i, j and
// loop-invariant integers v
struct Lin {
    bool ok = false;
    int64_t ci = 0, cj = 0, k = 0;
    std::map<std::string, int64_t> terms;
};
Lin lin_sum(const Lin& a, const Lin& b, int64_t sb) {
    Lin r;
    if (!a.ok || !b.ok) return r;
    r.ok = true;
    r.ci = a.ci + sb * b.ci;
    r.cj = a.cj + sb * b.cj;
    r.k = a.k + sb * b.k;
    r.terms = a.terms;
    for (const auto& t : b.terms) {
        r.terms[t.first] += sb * t.second;
        if (r.terms[t.first] == 0) r.terms.erase(t.first);
    }
    return r;
}
Lin lin_scale(const Lin& a, int64_t c) {
    Lin r = a;
    if (!r.ok) return r;
    r.ci *= c, r.cj *= c, r.k *= c;
    for (auto& t : r.terms) t.second *= c;
    if (c == 0) r.terms.clear();
    return r;
}

struct IterMember {
    Site sub;  // the loop as a site of its own: its arguments are its free variables
    Plan plan;
    bool near = false;
    std::string why;  // why the temporal form does not take it
    Lin wrow, wcol;   // where it writes the new state
    std::map<std::string, std::pair<Atom, Atom>> tile_pos;  // flat index variables of its loads
    std::map<std::string, std::pair<int64_t, int64_t>> rel;  // (row, column) offsets of those that move with the cell
    int64_t dmin[2] = {0, 0}, dmax[2] = {0, 0};             // offsets of the loads that move with the cell
    std::vector<Lin> fixed[2];                              // loads at a fixed row / column
    int radius = 0;
};

struct IterNode {
    enum Kind { Host, Member, If } kind = Host;
    const Stmt* s = nullptr;
    int member = -1;
    std::vector<IterNode> then_, else_;
};

struct IterPlan {
    std::string t, cur, next;
    const Stmt* loop = nullptr;
    const Stmt* alloc = nullptr;
    CType elem;
    std::vector<IterNode> nodes;
    std::vector<std::unique_ptr<IterMember>> members;  // stable: plans point into their sites
    std::vector<std::vector<int>> leaves;  // runs of member loops with nothing between them
    std::vector<int> leaf_radius;          // -1: no temporal form for the run
    std::map<std::string, CType> host_vars;
    std::vector<std::string> host_order;
    std::set<std::string> shape_vars;  // host variables holding the state's shape
};

constexpr int kIterTH = 32, kIterLW = 64, kIterBlock = 512, kIterMaxRadius = 16;
// output tile width: the 64-column LDS row less the halo on both sides
int iter_tile_w(int r);
int iter_steps_per_launch(int r) { return r == 0 ? 8 : std::max(1, 8 / r); }
int iter_tile_w(int r) { return kIterLW - 2 * iter_steps_per_launch(r) * r; }

bool is_iterate_site(const Site& site) {
    return site.loop.kind == Stmt::SLoop && site.loop.loop.role == Role::Iterate;
}

struct IterAnalyzer {
    const Site& site;
    IterPlan& ip;

    [[noreturn]] void fail(const std::string& w) const { throw Error("site " + site.name + ": iterate: " + w); }

    CType host_rhs_type(const RHS& r) {
        auto t = [&](const Atom& a) -> CType {
            if (a.kind == Atom::Int || a.kind == Atom::Bool || a.kind == Atom::Unit) return CType::scalar(CType::Int64);
            if (a.kind == Atom::Float) return CType::scalar(CType::Double);
            if (!a.is_var()) fail("string or vector atom in the step");
            auto h = ip.host_vars.find(a.var);
            if (h != ip.host_vars.end()) return h->second;
            bool found;
            CType at = arg_type(site, a.var, &found);
            if (!found) fail("the step reads " + a.var + " before assigning it");
            return at;
        };
        switch (r.kind) {
            case RHS::RAtom: return t(r.args[0]);
            case RHS::RBinOp:
                if (is_compare(r.bop)) return CType::scalar(CType::Bool);
                return CType::scalar(is_float_op(r.bop) ? CType::Double : CType::Int64);
            case RHS::RUnOp:
                switch (r.uop) {
                    case UnOp::CNot: return CType::scalar(CType::Bool);
                    case UnOp::CNeg: return t(r.args[0]);
                    case UnOp::CIntOf: return CType::scalar(CType::Int64);
                    case UnOp::CLog: case UnOp::CSin: case UnOp::CCos: case UnOp::CErf:
                        fail("log/sin/cos/erf in the step's scalar statements");
                    default: return CType::scalar(CType::Double);
                }
            case RHS::RTuple: case RHS::RShapeInit: return CType::scalar(CType::Tuple);
            case RHS::RArrayShape:
                if (!r.args[0].is_var() || r.args[0].var != ip.cur)
                    fail("the step takes the shape of an array other than the state");
                return CType::scalar(CType::Tuple);
            case RHS::RProj: case RHS::RShapeSize: case RHS::RShapeLast: return CType::scalar(CType::Int64);
            default: fail("a scalar statement of the step is not one the launcher evaluates");
        }
    }

    // The first `count` statements of ss (all of them by default).
    std::vector<IterNode> block(const std::vector<Stmt>& ss, size_t count = SIZE_MAX) {
        std::vector<IterNode> out;
        bool in_run = false;
        for (size_t si = 0; si < ss.size() && si < count; ++si) {
            const Stmt& s = ss[si];
            IterNode n;
            n.s = &s;
            if (s.kind == Stmt::SAssign) {
                if (s.var == ip.next) {
                    if (&s != ip.alloc) fail(ip.next + " assigned inside a branch");
                    in_run = false;
                    continue;
                }
                if (ip.host_vars.count(s.var)) fail(s.var + " assigned twice in the step");
                bool found;
                (void)arg_type(site, s.var, &found);
                if (found) fail("the step assigns the argument " + s.var);
                const CType t = host_rhs_type(s.rhs);
                ip.host_vars[s.var] = t;
                ip.host_order.push_back(s.var);
                if (s.rhs.kind == RHS::RArrayShape ||
                    (s.rhs.kind == RHS::RAtom && s.rhs.args[0].is_var() && ip.shape_vars.count(s.rhs.args[0].var)))
                    ip.shape_vars.insert(s.var);
                n.kind = IterNode::Host;
                in_run = false;
            } else if (s.kind == Stmt::SIf) {
                if (!s.a0.is_var() || !ip.host_vars.count(s.a0.var)) fail("a branch on something other than a step scalar");
                n.kind = IterNode::If;
                n.then_ = block(s.body);
                n.else_ = block(s.els);
                in_run = false;
            } else if (s.kind == Stmt::SLoop) {
                n.kind = IterNode::Member;
                n.member = member(s);
                if (!in_run) ip.leaves.emplace_back();
                ip.leaves.back().push_back(n.member);
                in_run = true;
            } else {
                fail("a statement of the step that is neither scalar, a branch nor a loop");
            }
            out.push_back(std::move(n));
        }
        return out;
    }

    int member(const Stmt& loop) {
        auto mp = std::make_unique<IterMember>();
        IterMember& m = *mp;
        m.sub.name = site.name + "_m" + std::to_string(ip.members.size());
        std::set<std::string> used, assigned;
        stmt_vars({loop}, used, assigned);
        for (const std::string& v : used) {
            if (assigned.count(v)) continue;
            if (v == ip.t) fail("a loop of the step reads the iteration counter");
            CType t;
            auto h = ip.host_vars.find(v);
            bool found;
            if (v == ip.next) {
                t.kind = CType::Array;
                t.elem = std::make_shared<CType>(ip.elem);
            } else if (h != ip.host_vars.end()) {
                t = h->second;
            } else {
                t = arg_type(site, v, &found);
                if (!found) fail("loop reads " + v + ", which is neither an argument nor a step scalar");
            }
            if (t.is_array() && v != ip.cur && v != ip.next) {
                // another array: read or written each step, fine for the in-order form
            }
            m.sub.args.push_back({v, t});
        }
        m.sub.types = site.types;
        for (const auto& h : ip.host_vars) m.sub.types[h.first] = h.second;
        m.sub.loop = loop;
        m.plan = make_plan(m.sub);
        if (m.plan.form != Plan::Map1D && m.plan.form != Plan::Map2D && m.plan.form != Plan::MapND)
            fail("a loop of the step is not a map (" + std::string(form_name(m.plan.form)) + ")");
        analyze_near(m);
        ip.members.push_back(std::move(mp));
        return (int)ip.members.size() - 1;
    }

    void analyze_near(IterMember& m) {
        const Plan& p = m.plan;
        auto no = [&](const std::string& w) { m.near = false, m.why = w; };
        if (p.form != Plan::Map2D) return no("not a 2-index map");
        for (const auto& a : p.arrays) {
            if (a.first == ip.cur) {
                if (a.second.written) return no("writes the old state");
            } else if (a.first == ip.next) {
                if (a.second.read) return no("reads the new state");
            } else {
                return no("touches the array " + a.first);
            }
        }
        if (!p.arrays.count(ip.next) || !p.arrays.at(ip.next).written) return no("does not write the new state");
        const std::vector<Stmt>& body = p.kloop->body;
        const std::string& I = p.iters[0];
        const std::string& J = p.iters[1];
        std::map<std::string, int> count;
        for_each_stmt(body, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign) ++count[s.var];
            if (s.kind == Stmt::SLoop)
                for (const auto& it : s.loop.iters) ++count[it];
        });
        std::set<std::string> free_ints;
        for (const auto& a : m.sub.args)
            if (!a.second.is_array() && a.second.is_int_like() && !count.count(a.first)) free_ints.insert(a.first);
        std::map<std::string, Lin> form;
        std::map<std::string, const Stmt*> def;
        auto atom_lin = [&](const Atom& a) -> Lin {
            Lin r;
            if (a.kind == Atom::Int) {
                r.ok = true, r.k = a.i;
            } else if (a.is_var()) {
                if (a.var == I) r.ok = true, r.ci = 1;
                else if (a.var == J) r.ok = true, r.cj = 1;
                else if (form.count(a.var)) r = form.at(a.var);
                else if (free_ints.count(a.var)) r.ok = true, r.terms[a.var] = 1;
            }
            return r;
        };
        for (const Stmt& s : body) {
            if (s.kind != Stmt::SAssign || count[s.var] != 1) continue;
            def[s.var] = &s;
            const RHS& r = s.rhs;
            Lin f;
            if (r.kind == RHS::RAtom) {
                f = atom_lin(r.args[0]);
            } else if (r.kind == RHS::RBinOp && (r.bop == BinOp::CAdd || r.bop == BinOp::CSub)) {
                f = lin_sum(atom_lin(r.args[0]), atom_lin(r.args[1]), r.bop == BinOp::CAdd ? 1 : -1);
            } else if (r.kind == RHS::RBinOp && r.bop == BinOp::CMul) {
                if (r.args[0].kind == Atom::Int) f = lin_scale(atom_lin(r.args[1]), r.args[0].i);
                else if (r.args[1].kind == Atom::Int) f = lin_scale(atom_lin(r.args[0]), r.args[1].i);
            }
            if (f.ok) form[s.var] = f;
        }
        // (row, column) of a flat index built as nd_to_flat((row, col), shape of the state)
        auto pos = [&](const Atom& f, Lin& r, Lin& c, std::pair<Atom, Atom>* atoms) -> bool {
            if (!f.is_var() || !def.count(f.var)) return false;
            const Stmt* d = def.at(f.var);
            if (d->rhs.kind != RHS::RNdToFlat) return false;
            const Atom& tup = d->rhs.args[0];
            const Atom& shp = d->rhs.args[1];
            if (!shp.is_var() || !ip.shape_vars.count(shp.var)) return false;
            if (!tup.is_var() || !def.count(tup.var)) return false;
            const Stmt* td = def.at(tup.var);
            if (td->rhs.kind != RHS::RTuple || td->rhs.args.size() != 2) return false;
            r = atom_lin(td->rhs.args[0]);
            c = atom_lin(td->rhs.args[1]);
            if (atoms) *atoms = {td->rhs.args[0], td->rhs.args[1]};
            return r.ok && c.ok;
        };
        int writes = 0;
        const Stmt* w = nullptr;
        for_each_stmt(body, [&](const Stmt& s) {
            if (s.kind == Stmt::SArrayWrite) ++writes;
        });
        for (const Stmt& s : body)
            if (s.kind == Stmt::SArrayWrite) w = &s;
        if (writes != 1 || !w) return no("not one top-level write of the new state");
        if (!pos(w->a1, m.wrow, m.wcol, nullptr)) return no("the write is not at nd_to_flat((row, col), shape)");
        if (!(m.wrow.ci == 1 && m.wrow.cj == 0 && m.wcol.ci == 0 && m.wcol.cj == 1))
            return no("the write is not at (i + a, j + b)");
        bool bad = false;
        std::string why;
        for_each_stmt(body, [&](const Stmt& s) {
            if (bad || s.kind != Stmt::SAssign || s.rhs.kind != RHS::RArrayLoad) return;
            Lin r, c;
            std::pair<Atom, Atom> rc;
            if (!pos(s.rhs.args[1], r, c, &rc)) {
                bad = true, why = "a load index is not nd_to_flat((row, col), shape)";
                return;
            }
            m.tile_pos[s.rhs.args[1].var] = rc;
            if (r.ci == 1 && c.cj == 1 && r.cj == 0 && c.ci == 0 && r.terms == m.wrow.terms && c.terms == m.wcol.terms)
                m.rel[s.rhs.args[1].var] = {r.k - m.wrow.k, c.k - m.wcol.k};
            const Lin* rcw[2] = {&r, &c};
            const Lin* wl[2] = {&m.wrow, &m.wcol};
            for (int d = 0; d < 2; ++d) {
                const Lin& x = *rcw[d];
                const int64_t own = d == 0 ? x.ci : x.cj, other = d == 0 ? x.cj : x.ci;
                if (other != 0 || (own != 0 && own != 1)) {
                    bad = true, why = "a load mixes the row and column iterators";
                    return;
                }
                if (own == 1) {
                    if (x.terms != wl[d]->terms) {
                        bad = true, why = "a load's offset from the cell is not a constant";
                        return;
                    }
                    const int64_t delta = x.k - wl[d]->k;
                    m.dmin[d] = std::min<int64_t>(m.dmin[d], delta);
                    m.dmax[d] = std::max<int64_t>(m.dmax[d], delta);
                } else {
                    m.fixed[d].push_back(x);
                }
            }
        });
        if (bad) return no(why);
        // the flat index variables of the loads serve only as load indices
        // (and possibly the write's): they print as tile offsets
        for_each_stmt(body, [&](const Stmt& s) {
            if (bad) return;
            auto check = [&](const Atom& a, bool ok_here) {
                if (a.is_var() && m.tile_pos.count(a.var) && !ok_here) bad = true;
            };
            if (s.kind == Stmt::SAssign) {
                for (size_t k = 0; k < s.rhs.args.size(); ++k)
                    check(s.rhs.args[k], s.rhs.kind == RHS::RArrayLoad && k == 1);
                if (m.tile_pos.count(s.var) && count[s.var] != 1) bad = true;
            } else if (s.kind == Stmt::SArrayWrite) {
                check(s.a0, false), check(s.a1, true), check(s.a2, false);
            } else if (s.kind == Stmt::SIf) {
                check(s.a0, false);
            }
        });
        if (bad) return no("a load's index variable is used as a value");
        int64_t r = 0;
        for (int d = 0; d < 2; ++d) r = std::max<int64_t>(r, std::max<int64_t>(-m.dmin[d], m.dmax[d]));
        if (r > kIterMaxRadius) return no("radius above " + std::to_string(kIterMaxRadius));
        m.radius = (int)r;
        m.near = true;
    }

    void run() {
        const Stmt& L = site.loop;
        if (L.loop.iters.size() != 1 || L.loop.bounds.size() != 1) fail("one counter expected");
        if (L.loop.exec != LoopSpec::Serial) fail("the iterate loop must be serial");
        if (L.loop.has_red) fail("a reduction on the iterate loop");
        ip.loop = &L;
        ip.t = L.loop.iters[0];
        if (!site.has_result || !site.result.second.is_array() || !site.result.second.elem)
            fail("the site returns the iterated array");
        ip.cur = site.result.first;
        bool found;
        const CType ct = arg_type(site, ip.cur, &found);
        if (!found || !ct.is_array() || !ct.elem) fail(ip.cur + " is not an array argument of the site");
        ip.elem = *ct.elem;
        (void)c_type(ip.elem);
        const std::vector<Stmt>& body = L.body;
        if (body.empty()) fail("an empty step");
        const Stmt& last = body.back();
        if (!(last.kind == Stmt::SAssign && last.var == ip.cur && last.rhs.kind == RHS::RAtom &&
              last.rhs.args[0].is_var()))
            fail("the step does not end with " + ip.cur + " = <new state>");
        ip.next = last.rhs.args[0].var;
        const std::vector<Stmt> step(body.begin(), body.end() - 1);  // for the variable checks only
        std::set<std::string> used, assigned;
        stmt_vars(step, used, assigned);
        if (used.count(ip.t)) fail("the step reads the iteration counter");
        if (assigned.count(ip.cur)) fail("the step assigns " + ip.cur + " before its end");
        int allocs = 0;
        for_each_stmt(step, [&](const Stmt& s) {
            if (s.kind == Stmt::SAssign && s.var == ip.next) ++allocs;
        });
        for (size_t k = 0; k + 1 < body.size(); ++k)
            if (body[k].kind == Stmt::SAssign && body[k].var == ip.next && body[k].rhs.kind == RHS::RArrayAlloc)
                ip.alloc = &body[k];
        if (allocs != 1 || !ip.alloc) fail("the step must allocate " + ip.next + " once, at its top level");
        ip.nodes = block(body, body.size() - 1);  // nodes point into the site's own loop
        const Atom& shp = ip.alloc->rhs.args[0];
        if (!shp.is_var() || !ip.shape_vars.count(shp.var))
            fail(ip.next + " is not allocated with the shape of " + ip.cur);
        for (const auto& leaf : ip.leaves) {
            int r = 0;
            bool ok = leaf.size() <= 64;
            for (int k : leaf) {
                ok = ok && ip.members[(size_t)k]->near;
                r = std::max(r, ip.members[(size_t)k]->radius);
            }
            ip.leaf_radius.push_back(ok ? r : -1);
        }
    }
};

IterPlan make_iter_plan(const Site& site) {
    IterPlan ip;
    IterAnalyzer{site, ip}.run();
    return ip;
}

std::string lin_host(const Lin& l, const std::string& prefix) {
    std::string s = "(int64_t)" + std::to_string(l.k) + "LL";
    for (const auto& t : l.terms)
        s += " + (int64_t)" + std::to_string(t.second) + "LL * " + prefix + sanitize(t.first);
    return "(" + s + ")";
}

// The temporal kernel of a run (hk_tk_<run>): tiles whose every computed
// cell lies in one loop's box (all but the grid's edge tiles of a split
// stencil) run that loop's body alone in the band form; the edge tiles,
// listed by the launcher, each cell finding its loop.
void emit_iter_temporal(std::ostringstream& o, const Site& site, const IterPlan& ip, size_t L) {
    const std::vector<int>& leaf = ip.leaves[L];
    const int R = ip.leaf_radius[L];
    const int K = iter_steps_per_launch(R), H = K * R;
    const int LW = kIterLW, TW = LW - 2 * H, LH = kIterTH + 2 * H;
    const int WAVES = kIterBlock / 64, RPW = (LH + WAVES - 1) / WAVES;
    const size_t M = leaf.size();
    const std::string T = c_type(ip.elem);
    const std::string tag = site.name + "_" + std::to_string(L);
    bool exp = false;
    for (int k : leaf) exp = exp || ip.members[(size_t)k]->plan.uses_exp;
    // loops that may own a whole tile: those whose loads all move with the
    // cell (a loop with loads at a fixed row or column, a clamped boundary,
    // is at most the halo wide there)
    std::vector<size_t> uni;
    for (size_t k = 0; k < M; ++k) {
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        if (m.fixed[0].empty() && m.fixed[1].empty()) uni.push_back(k);
    }
    o << "// temporal form of " << site.name << " run " << L << ": radius " << R << ", " << K
      << " steps per launch over " << kIterTH << " x " << TW << " tiles with a " << H << "-cell halo (LDS rows of "
      << LW << ": a lane per column, a wave per row)\n";
    o << "struct hk_tb_" << tag << " {\n    int64_t rlo[" << M << "], rhi[" << M << "], clo[" << M << "], chi[" << M
      << "], wr[" << M << "], wc[" << M << "];\n};\n";
    // the kernels' arguments beyond the arrays: the boxes and every loop's
    // parameters, one block the launcher uploads once per call
    o << "struct hk_ta_" << tag << " {\n    hk_tb_" << tag << " tb;\n";
    for (size_t k = 0; k < M; ++k) o << "    hk_m_" << ip.members[(size_t)leaf[k]]->sub.name << " P" << k << ";\n";
    o << "};\n";
    // one member's body for the cell (hk_r, hk_c), its value into hk_val
    auto body = [&](size_t k, int ind) {
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        const std::string ks = std::to_string(k);
        const std::string pad((size_t)ind * 4, ' ');
        for (const Param& q : kernel_params(m.sub, m.plan))
            o << pad << "const " << q.ctype << " " << q.name << " = hk_A->P" << k << "." << q.name << ";\n"
              << pad << "(void)" << q.name << ";\n";
        Printer pr(m.sub, m.plan, o);
        pr.tile_array = ip.cur;
        pr.out_array = ip.next;
        pr.tile_pos = &m.tile_pos;
        pr.declare_locals(ind);
        o << pad << "const int64_t " << sanitize(m.plan.iters[0]) << " = hk_r - hk_tb.wr[" << ks << "];\n";
        o << pad << "const int64_t " << sanitize(m.plan.iters[1]) << " = hk_c - hk_tb.wc[" << ks << "];\n";
        std::set<std::string> assigned;
        pr.stmts(m.plan.kloop->body, ind, assigned);
    };
    // kind 0: the band kernel over every tile (a tile no loop owns whole
    // returns); kind 1: the edge kernel over the launcher's list of those
    for (int kind = 0; kind < 2; ++kind) {
    o << "__global__ __launch_bounds__(" << kIterBlock << ") void hk_tk" << (kind ? "e" : "") << "_" << tag
      << "(const " << T << "* __restrict__ hk_in, " << T
      << "* __restrict__ hk_out, int64_t hk_g0, int64_t hk_g1, int hk_steps, const hk_ta_" << tag
      << "* __restrict__ hk_A";
    if (kind) {
        o << ", const uint32_t* __restrict__ hk_tiles, uint32_t hk_ntx) {\n";
        o << "    const uint32_t hk_tile = hk_tiles[blockIdx.x];\n";
        o << "    const int64_t hk_bx = hk_tile % hk_ntx, hk_by = hk_tile / hk_ntx;\n";
    } else {
        o << ") {\n";
        o << "    const int64_t hk_bx = blockIdx.x, hk_by = blockIdx.y;\n";
    }
    o << "    const hk_tb_" << tag << "& hk_tb = hk_A->tb;\n";
    o << "    constexpr int hk_TH = " << kIterTH << ", hk_TW = " << TW << ", hk_H = " << H << ", hk_RR = " << R << ";\n";
    o << "    constexpr int hk_LH = " << LH << ", hk_LW = " << LW << ", hk_NC = hk_LH * hk_LW, hk_WAVES = " << WAVES
      << ", hk_RPW = " << RPW << ";\n";
    o << "    const int64_t hk_R0 = hk_by * hk_TH - hk_H, hk_C0 = hk_bx * hk_TW - hk_H;\n";
    // the loop whose box holds every cell the workgroup computes (its first
    // step's region), if one does (the launcher's edge list applies the
    // same test)
    o << "    int hk_uni = -1;\n";
    o << "    {\n";
    o << "        const int64_t hk_a0 = max(hk_R0 + hk_RR, (int64_t)0), hk_a1 = min(hk_R0 + hk_LH - hk_RR, hk_g0);\n";
    o << "        const int64_t hk_b0 = max(hk_C0 + hk_RR, (int64_t)0), hk_b1 = min(hk_C0 + hk_LW - hk_RR, hk_g1);\n";
    bool first_u = true;
    for (size_t k : uni) {
        const std::string ks = std::to_string(k);
        o << "        " << (first_u ? "" : "else ") << "if (hk_a0 >= hk_tb.rlo[" << ks << "] && hk_a1 <= hk_tb.rhi[" << ks
          << "] && hk_b0 >= hk_tb.clo[" << ks << "] && hk_b1 <= hk_tb.chi[" << ks << "]) hk_uni = " << ks << ";\n";
        first_u = false;
    }
    o << "    }\n";
    if (kind) o << "    (void)hk_uni;\n";
    else o << "    if (hk_uni < 0) return;\n";
    o << "    __shared__ " << T << " hk_tile0[hk_NC], hk_tile1[hk_NC];\n";
    if (exp)
        o << "    __shared__ uint64_t hk_exptab[256];\n"
             "    for (int hk_t = threadIdx.x; hk_t < 256; hk_t += blockDim.x) hk_exptab[hk_t] = hyd_exp_tab_dev[hk_t];\n";
    o << "    const int hk_lane = threadIdx.x & 63, hk_wave = threadIdx.x >> 6;\n";
    o << "    const int64_t hk_c = hk_C0 + hk_lane;\n";
    o << "    const bool hk_cin = hk_c >= 0 && hk_c < hk_g1;\n";
    o << "#pragma unroll\n";
    o << "    for (int hk_k = 0; hk_k < hk_RPW; ++hk_k) {\n";
    o << "        const int hk_lr = hk_wave + hk_k * hk_WAVES;\n";
    o << "        const int64_t hk_r = hk_R0 + hk_lr;\n";
    o << "        if (hk_lr < hk_LH && hk_cin && hk_r >= 0 && hk_r < hk_g0) hk_tile0[hk_lr * hk_LW + hk_lane] = hk_in[hk_r * hk_g1 + hk_c];\n";
    o << "    }\n";
    o << "    __syncthreads();\n";
    o << "    " << T << "* hk_src = hk_tile0;\n    " << T << "* hk_dst = hk_tile1;\n";
    o << "    for (int hk_s = 1; hk_s <= hk_steps; ++hk_s) {\n";
    o << "        const int64_t hk_vr0 = max(hk_R0 + (int64_t)hk_s * hk_RR, (int64_t)0);\n";
    o << "        const int64_t hk_vr1 = min(hk_R0 + hk_LH - (int64_t)hk_s * hk_RR, hk_g0);\n";
    o << "        const bool hk_cok = hk_c >= max(hk_C0 + (int64_t)hk_s * hk_RR, (int64_t)0) &&\n";
    o << "                            hk_c < min(hk_C0 + hk_LW - (int64_t)hk_s * hk_RR, hk_g1);\n";
    if (kind) {
    o << "        {\n";
    o << "            for (int hk_lr = hk_wave; hk_lr < hk_LH; hk_lr += hk_WAVES) {\n";
    o << "                const int64_t hk_r = hk_R0 + hk_lr;\n";
    o << "                if (!hk_cok || hk_r < hk_vr0 || hk_r >= hk_vr1) continue;\n";
    o << "                const int hk_q = hk_lr * hk_LW + hk_lane;\n";
    o << "                " << T << " hk_val = hk_src[hk_q];\n";
    for (size_t k = 0; k < M; ++k) {
        const std::string ks = std::to_string(k);
        o << "                " << (k ? "else " : "") << "if (hk_r >= hk_tb.rlo[" << ks << "] && hk_r < hk_tb.rhi[" << ks
          << "] && hk_c >= hk_tb.clo[" << ks << "] && hk_c < hk_tb.chi[" << ks << "]) {\n";
        body(k, 5);
        o << "                }\n";
    }
    o << "                hk_dst[hk_q] = hk_val;\n";
    o << "            }\n";
    } else {
    if (uni.empty()) o << "        {\n";
    // band form: wave w owns rows [w * RPW, (w + 1) * RPW) of the tile; per
    // step it reads its rows and R above and below into registers once
    // (rows clamped into the tile: the clamped ones feed only cells that
    // are not stored), takes the column neighbours by lane shifts, and
    // computes every row of the band
    for (size_t u = 0; u < uni.size(); ++u) {
        const size_t k = uni[u];
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        std::set<int64_t> dcs = {0};
        for (const auto& rp : m.rel) dcs.insert(rp.second.second);
        o << "        " << (u ? "} else " : "") << "if (hk_uni == " << k << ") {\n";
        o << "            const int hk_lr0 = hk_wave * hk_RPW;\n";
        o << "            " << T << " hk_x0[hk_RPW + 2 * hk_RR];\n";
        o << "#pragma unroll\n";
        o << "            for (int hk_q = 0; hk_q < hk_RPW + 2 * hk_RR; ++hk_q)\n";
        o << "                hk_x0[hk_q] = hk_src[min(max(hk_lr0 - hk_RR + hk_q, 0), hk_LH - 1) * hk_LW + hk_lane];\n";
        for (int64_t dc : dcs) {
            if (dc == 0) continue;
            o << "            " << T << " hk_x" << dc_name(dc) << "[hk_RPW + 2 * hk_RR];\n";
            o << "#pragma unroll\n";
            o << "            for (int hk_q = 0; hk_q < hk_RPW + 2 * hk_RR; ++hk_q) hk_x" << dc_name(dc)
              << "[hk_q] = hydk_lane_shift<" << dc << ">(hk_x0[hk_q]);\n";
        }
        o << "#pragma unroll\n";
        o << "            for (int hk_k = 0; hk_k < hk_RPW; ++hk_k) {\n";
        o << "                const int hk_lr = hk_lr0 + hk_k;\n";
        o << "                const int64_t hk_r = hk_R0 + hk_lr;\n";
        o << "                " << T << " hk_val = hk_x0[hk_k + hk_RR];\n";
        o << "                {\n";
        {
            const std::string ks = std::to_string(k);
            const std::string pad((size_t)5 * 4, ' ');
            for (const Param& q : kernel_params(m.sub, m.plan))
                o << pad << "const " << q.ctype << " " << q.name << " = hk_A->P" << k << "." << q.name << ";\n"
                  << pad << "(void)" << q.name << ";\n";
            Printer pr(m.sub, m.plan, o);
            pr.tile_array = ip.cur;
            pr.out_array = ip.next;
            pr.tile_pos = &m.tile_pos;
            pr.reg_pos = &m.rel;
            pr.reg_rr = R;
            pr.declare_locals(5);
            o << pad << "const int64_t " << sanitize(m.plan.iters[0]) << " = hk_r - hk_tb.wr[" << ks << "];\n";
            o << pad << "const int64_t " << sanitize(m.plan.iters[1]) << " = hk_c - hk_tb.wc[" << ks << "];\n";
            std::set<std::string> assigned;
            pr.stmts(m.plan.kloop->body, 5, assigned);
        }
        o << "                }\n";
        o << "                if (hk_cok && hk_lr < hk_LH && hk_r >= hk_vr0 && hk_r < hk_vr1)\n";
        o << "                    hk_dst[hk_lr * hk_LW + hk_lane] = hk_val;\n";
        o << "            }\n";
    }
    }
    o << "        }\n";
    o << "        __syncthreads();\n";
    o << "        " << T << "* hk_t = hk_src;\n        hk_src = hk_dst;\n        hk_dst = hk_t;\n";
    o << "    }\n";
    o << "    if (hk_lane >= hk_H && hk_lane < hk_H + hk_TW && hk_cin)\n";
    o << "        for (int hk_lr = hk_H + hk_wave; hk_lr < hk_H + hk_TH; hk_lr += hk_WAVES) {\n";
    o << "            const int64_t hk_r = hk_R0 + hk_lr;\n";
    o << "            if (hk_r < hk_g0) hk_out[hk_r * hk_g1 + hk_c] = hk_src[hk_lr * hk_LW + hk_lane];\n";
    o << "        }\n";
    o << "}\n\n";
    }
}

// The launcher's attempt at the temporal form for run L (inside the
// launcher, after the step scalars): true when it ran every step.
void emit_iter_try(std::ostringstream& o, const IterPlan& ip, size_t L, const std::string& name) {
    const std::vector<int>& leaf = ip.leaves[L];
    const int R = ip.leaf_radius[L];
    const int K = iter_steps_per_launch(R), TW = iter_tile_w(R), H = K * R, LH = kIterTH + 2 * H;
    const size_t M = leaf.size();
    const std::string tag = name + "_" + std::to_string(L);
    const std::string T = c_type(ip.elem);
    o << "    if (!hk_done && hk_nact == " << M;
    for (size_t k = 0; k < M; ++k) o << " && hk_act[" << k << "] == " << leaf[k];
    o << ") {\n";
    o << "        " << sanitize(ip.cur) << " = hk_cur_in;\n        " << sanitize(ip.next) << " = hk_buf[0];\n";
    for (size_t k = 0; k < M; ++k) {
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        std::string args;
        for (const auto& a : m.sub.args) args += (args.empty() ? "" : ", ") + sanitize(a.first);
        o << "        hk_x_" << m.sub.name << " hk_x" << k << ";\n";
        o << "        const hk_m_" << m.sub.name << " hk_P" << k << " = hk_prep_" << m.sub.name << "(" << args
          << (args.empty() ? "" : ", ") << "&hk_x" << k << ");\n";
    }
    o << "        const int64_t hk_g0 = hk_shape.elems[0], hk_g1 = hk_shape.ndims == 2 ? hk_shape.elems[1] : 0;\n";
    o << "        bool hk_ok = hk_shape.ndims == 2 && hk_g0 > 0 && hk_g1 > 0 && (hk_g0 + " << kIterTH - 1 << ") / "
      << kIterTH << " <= 65535 && (hk_g1 + " << TW - 1 << ") / " << TW << " <= 0x7fffffff;\n";
    o << "        hk_tb_" << tag << " hk_tb;\n        memset(&hk_tb, 0, sizeof hk_tb);\n";
    o << "        int64_t hk_area = 0;\n";
    for (size_t k = 0; k < M; ++k) {
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        const std::string ks = std::to_string(k), P = "hk_P" + ks + ".";
        o << "        if (hk_x" << ks << ".hk_live) {\n";
        o << "            hk_tb.wr[" << ks << "] = " << lin_host(m.wrow, P) << ";\n";
        o << "            hk_tb.wc[" << ks << "] = " << lin_host(m.wcol, P) << ";\n";
        o << "            hk_tb.rlo[" << ks << "] = " << P << "hk_lo0 + hk_tb.wr[" << ks << "];\n";
        o << "            hk_tb.rhi[" << ks << "] = hk_tb.rlo[" << ks << "] + " << P << "hk_n0;\n";
        o << "            hk_tb.clo[" << ks << "] = " << P << "hk_lo1 + hk_tb.wc[" << ks << "];\n";
        o << "            hk_tb.chi[" << ks << "] = hk_tb.clo[" << ks << "] + " << P << "hk_n1;\n";
        o << "            hk_ok = hk_ok && hk_tb.rlo[" << ks << "] >= 0 && hk_tb.rhi[" << ks << "] <= hk_g0 && hk_tb.clo[" << ks
          << "] >= 0 && hk_tb.chi[" << ks << "] <= hk_g1;\n";
        // loads that move with the cell stay inside the grid
        o << "            hk_ok = hk_ok && hk_tb.rlo[" << ks << "] + (" << m.dmin[0] << ") >= 0 && hk_tb.rhi[" << ks << "] - 1 + ("
          << m.dmax[0] << ") < hk_g0 && hk_tb.clo[" << ks << "] + (" << m.dmin[1] << ") >= 0 && hk_tb.chi[" << ks
          << "] - 1 + (" << m.dmax[1] << ") < hk_g1;\n";
        // loads at a fixed row / column: inside the grid and within the radius of every cell
        for (int d = 0; d < 2; ++d)
            for (const Lin& f : m.fixed[d]) {
                const std::string lo = d == 0 ? "hk_tb.rlo[" + ks + "]" : "hk_tb.clo[" + ks + "]";
                const std::string hi = d == 0 ? "hk_tb.rhi[" + ks + "]" : "hk_tb.chi[" + ks + "]";
                const std::string g = d == 0 ? "hk_g0" : "hk_g1";
                o << "            { const int64_t hk_v = " << lin_host(f, P) << ";\n";
                o << "              hk_ok = hk_ok && hk_v >= 0 && hk_v < " << g << " && hk_v - " << lo << " <= " << R
                  << " && " << lo << " - hk_v <= " << R << " && hk_v - (" << hi << " - 1) <= " << R << " && (" << hi
                  << " - 1) - hk_v <= " << R << "; }\n";
            }
        o << "            if (hk_tb.rhi[" << ks << "] > hk_tb.rlo[" << ks << "] && hk_tb.chi[" << ks << "] > hk_tb.clo[" << ks
          << "]) hk_area += (hk_tb.rhi[" << ks << "] - hk_tb.rlo[" << ks << "]) * (hk_tb.chi[" << ks << "] - hk_tb.clo[" << ks
          << "]);\n";
        o << "        }\n";
    }
    // the boxes tile the grid: disjoint, and their areas add up to it
    o << "        for (int hk_a = 0; hk_ok && hk_a < " << M << "; ++hk_a)\n";
    o << "            for (int hk_b = hk_a + 1; hk_ok && hk_b < " << M << "; ++hk_b)\n";
    o << "                hk_ok = !(hk_tb.rlo[hk_a] < hk_tb.rhi[hk_b] && hk_tb.rlo[hk_b] < hk_tb.rhi[hk_a] &&\n";
    o << "                          hk_tb.clo[hk_a] < hk_tb.chi[hk_b] && hk_tb.clo[hk_b] < hk_tb.chi[hk_a] &&\n";
    o << "                          hk_tb.rhi[hk_a] > hk_tb.rlo[hk_a] && hk_tb.chi[hk_a] > hk_tb.clo[hk_a] &&\n";
    o << "                          hk_tb.rhi[hk_b] > hk_tb.rlo[hk_b] && hk_tb.chi[hk_b] > hk_tb.clo[hk_b]);\n";
    o << "        hk_ok = hk_ok && hk_g0 <= 0x7fffffffLL && hk_g1 <= 0x7fffffffLL && hk_area == hk_g0 * hk_g1;\n";
    o << "        const dim3 hk_grid(hk_ok ? (unsigned)((hk_g1 + " << TW - 1 << ") / " << TW << ") : 0u, hk_ok ? (unsigned)((hk_g0 + "
      << kIterTH - 1 << ") / " << kIterTH << ") : 0u);\n";
    o << "        std::vector<uint32_t> hk_edge;\n";
    o << "        if (hk_ok) {\n";
    // the tiles no loop owns whole (the kernel's own test, once per call)
    o << "            for (uint32_t hk_by = 0; hk_by < hk_grid.y; ++hk_by)\n";
    o << "                for (uint32_t hk_bx = 0; hk_bx < hk_grid.x; ++hk_bx) {\n";
    o << "                    const int64_t hk_R0 = (int64_t)hk_by * " << kIterTH << " - " << H << ", hk_C0 = (int64_t)hk_bx * "
      << TW << " - " << H << ";\n";
    o << "                    const int64_t hk_a0 = std::max<int64_t>(hk_R0 + " << R << ", 0), hk_a1 = std::min<int64_t>(hk_R0 + "
      << LH - R << ", hk_g0);\n";
    o << "                    const int64_t hk_b0 = std::max<int64_t>(hk_C0 + " << R << ", 0), hk_b1 = std::min<int64_t>(hk_C0 + "
      << kIterLW - R << ", hk_g1);\n";
    o << "                    bool hk_whole = false;\n";
    for (size_t k = 0; k < M; ++k) {
        const IterMember& m = *ip.members[(size_t)leaf[k]];
        if (!(m.fixed[0].empty() && m.fixed[1].empty())) continue;
        const std::string ks = std::to_string(k);
        o << "                    hk_whole = hk_whole || (hk_a0 >= hk_tb.rlo[" << ks << "] && hk_a1 <= hk_tb.rhi[" << ks
          << "] && hk_b0 >= hk_tb.clo[" << ks << "] && hk_b1 <= hk_tb.chi[" << ks << "]);\n";
    }
    o << "                    if (!hk_whole) hk_edge.push_back(hk_by * hk_grid.x + hk_bx);\n";
    o << "                }\n";
    o << "            if (hk_edge.size() > 0x7fffffffu) hk_ok = false;\n";
    o << "        }\n";
    o << "        if (hk_ok) {\n";
    // the argument block, then the edge list, in one scratch buffer
    o << "            hk_ta_" << tag << " hk_ta;\n";
    o << "            hk_ta.tb = hk_tb;\n";
    for (size_t k = 0; k < M; ++k) o << "            hk_ta.P" << k << " = hk_P" << k << ";\n";
    o << "            const size_t hk_ta_bytes = (sizeof hk_ta + 255) / 256 * 256;\n";
    o << "            std::vector<unsigned char> hk_up(hk_ta_bytes + hk_edge.size() * 4);\n";
    o << "            memcpy(hk_up.data(), &hk_ta, sizeof hk_ta);\n";
    o << "            if (!hk_edge.empty()) memcpy(hk_up.data() + hk_ta_bytes, hk_edge.data(), hk_edge.size() * 4);\n";
    o << "            unsigned char* hk_up_d = (unsigned char*)hydk_scratch(3, (int64_t)hk_up.size(), hk_s);\n";
    o << "            if (hipMemcpyAsync(hk_up_d, hk_up.data(), hk_up.size(), hipMemcpyHostToDevice, hk_s) != hipSuccess ||\n";
    o << "                hipStreamSynchronize(hk_s) != hipSuccess)\n";
    o << "                hydk_fail(hk_site, \"uploading the kernel arguments failed\");\n";
    o << "            const hk_ta_" << tag << "* hk_ta_d = (const hk_ta_" << tag << "*)hk_up_d;\n";
    o << "            const uint32_t* hk_edge_d = (const uint32_t*)(hk_up_d + hk_ta_bytes);\n";
    o << "            for (int64_t hk_left = hk_steps; hk_left > 0;) {\n";
    o << "                const int hk_k = (int)std::min<int64_t>(" << K << ", hk_left);\n";
    o << "                if (!hk_buf[hk_w]) hk_buf[hk_w] = hk_alloc(hk_shape, hk_site);\n";
    for (int kind = 0; kind < 2; ++kind) {
        if (kind)
            o << "                if (!hk_edge.empty()) hk_tke_" << tag << "<<<(unsigned)hk_edge.size(), " << kIterBlock
              << ", 0, hk_s>>>(";
        else
            o << "                hk_tk_" << tag << "<<<hk_grid, " << kIterBlock << ", 0, hk_s>>>(";
        o << "(const " << T << "*)hydk_device_data(hk_src), (" << T << "*)hydk_device_data(hk_buf[hk_w]), hk_g0, hk_g1, hk_k, hk_ta_d";
        if (kind) o << ", hk_edge_d, hk_grid.x";
        o << ");\n";
        o << "                hydk_check_launch(hk_site);\n";
    }
    o << "                hydk_device_written(hk_buf[hk_w]);\n";
    o << "                hk_src = hk_buf[hk_w];\n                hk_w ^= 1;\n                hk_left -= hk_k;\n";
    o << "            }\n";
    o << "            hk_done = true;\n";
    o << "        }\n";
    o << "    }\n";
}

void emit_iter_nodes(std::ostringstream& o, Printer& hostp, const std::vector<IterNode>& nodes, int ind) {
    for (const IterNode& n : nodes) {
        const std::string pad((size_t)ind * 4, ' ');
        if (n.kind == IterNode::Host) {
            const CType& t = hostp.p.locals.at(n.s->var);
            o << pad << sanitize(n.s->var) << " = (" << c_type(t) << ")" << hostp.rhs(n.s->rhs, t) << ";\n";
        } else if (n.kind == IterNode::Member) {
            o << pad << "hk_act[hk_nact++] = " << n.member << ";\n";
        } else {
            o << pad << "if (" << sanitize(n.s->a0.var) << ") {\n";
            emit_iter_nodes(o, hostp, n.then_, ind + 1);
            o << pad << "} else {\n";
            emit_iter_nodes(o, hostp, n.else_, ind + 1);
            o << pad << "}\n";
        }
    }
}

void emit_iterate_site(std::ostringstream& o, const Site& site, const Options& opt) {
    const IterPlan ip = make_iter_plan(site);
    o << "// ---- " << site.name << ": iterate of " << ip.members.size() << " loop(s)";
    for (size_t L = 0; L < ip.leaves.size(); ++L)
        o << "; run " << L << " " << (ip.leaf_radius[L] >= 0 ? "temporal" : "in order");
    o << "\n";
    for (const auto& m : ip.members) emit_site(o, m->sub, opt);
    for (size_t L = 0; L < ip.leaves.size(); ++L)
        if (ip.leaf_radius[L] >= 0) emit_iter_temporal(o, site, ip, L);
    const std::string T = c_type(ip.elem);
    std::string lparams;
    for (const auto& a : site.args) {
        lparams += lparams.empty() ? "" : ", ";
        const std::string nm = a.first == ip.cur ? "hk_cur_in" : sanitize(a.first);
        lparams += a.second.is_array() ? "hyd_array_t* " + nm : std::string(c_type(a.second)) + " " + nm;
    }
    // host evaluation of the step's scalars: a printer over a plan whose
    // locals are those scalars
    Plan hp;
    for (const auto& h : ip.host_vars) hp.locals[h.first] = h.second;
    Printer hostp(site, hp, o);
    hostp.host = true;
    o << "static hyd_array_t* hk_alloc_" << site.name << "(hyd_tuple_t hk_shape, const char* hk_site) {\n";
    o << "    hyd_array_t* hk_a = hyd_array_alloc_device(hk_shape, sizeof(" << T << "));\n";
    o << "    if (!hk_a) hydk_fail(hk_site, \"no memory for the iterate's state\");\n";
    o << "    return hk_a;\n}\n";
    // which form the last call took (0: no step, 1: in order, 2: temporal),
    // for tests and profiles
    o << "static int hk_form_" << site.name << " = 0;\n";
    o << "extern \"C\" int " << site.name << "__form(void) { return hk_form_" << site.name << "; }\n";
    o << "extern \"C\" hyd_array_t* " << site.name << "(" << lparams << ") {\n";
    o << "    const char* hk_site = \"" << site.name << "\";\n";
    o << "    hipStream_t hk_s = hydk_stream();\n";
    o << "    const int64_t hk_steps = " << hostp.index_expr(ip.loop->loop.bounds[0]) << ";\n";
    o << "    hk_form_" << site.name << " = 0;\n";
    o << "    if (hk_steps <= 0) return hk_cur_in;\n";
    o << "    auto hk_alloc = [](hyd_tuple_t hk_sh, const char* hk_nm) { return hk_alloc_" << site.name
      << "(hk_sh, hk_nm); };\n";
    o << "    hyd_array_t* " << sanitize(ip.cur) << " = hk_cur_in;\n";
    o << "    hyd_array_t* " << sanitize(ip.next) << " = nullptr;\n";
    for (const std::string& v : ip.host_order) {
        const CType& t = ip.host_vars.at(v);
        o << "    " << c_type(t) << " " << sanitize(v) << (t.kind == CType::Tuple ? " = hydk_tuple(0);\n" : " = 0;\n");
    }
    o << "    int hk_act[" << std::max<size_t>(1, ip.members.size()) << "];\n    int hk_nact = 0;\n";
    emit_iter_nodes(o, hostp, ip.nodes, 1);
    o << "    const hyd_tuple_t hk_shape = hk_cur_in->shape;\n";
    o << "    hyd_array_t* hk_buf[2] = {hk_alloc(hk_shape, hk_site), nullptr};\n";
    o << "    hyd_array_t* hk_src = hk_cur_in;\n";
    o << "    int hk_w = 0;\n";
    o << "    bool hk_done = false;\n";
    for (size_t L = 0; L < ip.leaves.size(); ++L)
        if (ip.leaf_radius[L] >= 0) emit_iter_try(o, ip, L, site.name);
    // in order, step by step, each step into a zeroed array
    o << "    hk_form_" << site.name << " = hk_done ? 2 : 1;\n";
    o << "    if (!hk_done) {\n";
    o << "        const int64_t hk_bytes = hyd_shape_size(hk_shape) * (int64_t)sizeof(" << T << ");\n";
    o << "        for (int64_t hk_t = 0; hk_t < hk_steps; ++hk_t) {\n";
    o << "            if (!hk_buf[hk_w]) hk_buf[hk_w] = hk_alloc(hk_shape, hk_site);\n";
    o << "            if (hk_bytes > 0 && hipMemsetAsync(hydk_device_data(hk_buf[hk_w]), 0, (size_t)hk_bytes, hk_s) != hipSuccess)\n";
    o << "                hydk_fail(hk_site, \"clearing the new state failed\");\n";
    o << "            hydk_device_written(hk_buf[hk_w]);\n";
    o << "            " << sanitize(ip.cur) << " = hk_src;\n";
    o << "            " << sanitize(ip.next) << " = hk_buf[hk_w];\n";
    o << "            for (int hk_a = 0; hk_a < hk_nact; ++hk_a) {\n";
    o << "                switch (hk_act[hk_a]) {\n";
    for (size_t k = 0; k < ip.members.size(); ++k) {
        const IterMember& m = *ip.members[k];
        std::string args;
        for (const auto& a : m.sub.args) args += (args.empty() ? "" : ", ") + sanitize(a.first);
        o << "                    case " << k << ": (void)" << m.sub.name << "(" << args << "); break;\n";
    }
    o << "                    default: break;\n";
    o << "                }\n";
    o << "            }\n";
    o << "            hk_src = hk_buf[hk_w];\n";
    o << "            hk_w ^= 1;\n";
    o << "        }\n";
    o << "    }\n";
    o << "    (void)" << sanitize(ip.cur) << ";\n    (void)" << sanitize(ip.next) << ";\n";
    o << "    for (int hk_b = 0; hk_b < 2; ++hk_b)\n";
    o << "        if (hk_buf[hk_b] && hk_buf[hk_b] != hk_src) hyd_array_free(hk_buf[hk_b]);\n";
    o << "    return hk_src;\n}\n\n";
}

std::string describe_iterate(const Site& site) {
    const IterPlan ip = make_iter_plan(site);
    std::ostringstream o;
    o << "form=iterate state=" << ip.cur << " next=" << ip.next << " loops=" << ip.members.size() << " runs=";
    for (size_t L = 0; L < ip.leaves.size(); ++L) {
        o << (L ? "," : "") << ip.leaves[L].size();
        if (ip.leaf_radius[L] >= 0)
            o << "(temporal radius=" << ip.leaf_radius[L] << " steps=" << iter_steps_per_launch(ip.leaf_radius[L])
              << " tile=" << kIterTH << "x" << iter_tile_w(ip.leaf_radius[L]) << ")";
        else
            o << "(in order)";
    }
    for (size_t k = 0; k < ip.members.size(); ++k)
        if (!ip.members[k]->near) o << " loop" << k << "=" << ip.members[k]->why;
    return o.str();
}

}  // namespace

std::string describe_plan(const Site& site, const Options& opt) {
    if (is_iterate_site(site)) return describe_iterate(site);
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
        if (is_iterate_site(s))
            emit_iterate_site(o, s, opt);
        else
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
        const std::string ret =
            !s.has_result ? "int64_t" : s.result.second.is_array() ? "hyd_array_t*" : c_type(s.result.second);
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

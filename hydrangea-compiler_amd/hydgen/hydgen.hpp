// hydgen.hpp - the HIP kernel generator of the MI355X backend.
//
// Input: offload sites chosen by Language.Hydrangea.CodegenHIP, one per line,
// each the derived `Show` text of a HipSite record whose fields hold the
// reference's own CFG values (src/Language/Hydrangea/CFG.hs, CFGCore.hs: every
// type there derives Show).  Output: one HIP source with a kernel (or two) and
// an extern "C" launcher per site, the launcher's name and parameters being
// those of the RCall that replaced the loop in the C the C backend prints.
//
// Stages: parse the Show text (show_parse.cpp) -> CFG mirror (cfg.cpp) ->
// GPU-schedule pass and HIP text (codegen.cpp: kernel form, grid, workgroup,
// elements per thread, stream loads, reduction tree, scatter merge).
#pragma once

#include <stdint.h>

#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hydgen {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---- derived-Show values ------------------------------------------------------
struct Value;
using VPtr = std::shared_ptr<const Value>;

struct Value {
    enum Kind { Con, Record, List, Tuple, Str, Num } kind = Con;
    std::string text;  // Con/Record: constructor; Str: the string; Num: literal text
    std::vector<VPtr> items;                            // Con arguments, List/Tuple items
    std::vector<std::pair<std::string, VPtr>> fields;  // Record fields in order

    const Value& field(const std::string& name) const;
    bool is_con(const char* name) const { return (kind == Con || kind == Record) && text == name; }
};

// Parses one value; throws Error with the offset of a syntax error.
VPtr parse_show(const std::string& text);
// Parses every non-empty line of a file's text.
std::vector<VPtr> parse_show_lines(const std::string& text);

// ---- CFG mirror (CFGCore.hs, CFG.hs) ----------------------------------------------
enum class BinOp {
    CAdd, CSub, CMul, CDiv, CMod, CEq, CNeq, CLt, CLe, CGt, CGe, CAnd, COr,
    CAddF, CSubF, CMulF, CDivF, CEqF, CNeqF, CLtF, CLeF, CGtF, CGeF
};
enum class UnOp { CNot, CNeg, CSqrt, CExpF, CLog, CSin, CCos, CAbsF, CFloorF, CCeilF, CErf, CFloatOf, CIntOf };

struct Atom {
    enum Kind { Var, Int, Float, Bool, Unit, String, VecVar } kind = Unit;
    std::string var;  // Var / VecVar / String
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    bool is_var() const { return kind == Var; }
};

struct CType {
    enum Kind { Int64, Double, Bool, Unit, Tuple, Array, Pair, Record, Unknown } kind = Unknown;
    std::shared_ptr<CType> elem;  // Array
    static CType scalar(Kind k) {
        CType t;
        t.kind = k;
        return t;
    }
    bool is_array() const { return kind == Array; }
    bool is_float() const { return kind == Double; }
    bool is_int_like() const { return kind == Int64 || kind == Bool; }
    std::string show() const;
};

struct RHS {
    enum Kind {
        RAtom, RBinOp, RUnOp, RTuple, RProj, RRecord, RRecordProj, RPairMake, RPairFst, RPairSnd,
        RArrayAlloc, RArrayCopy, RArrayLoad, RArrayShape, RShapeSize, RShapeInit, RShapeLast,
        RFlatToNd, RNdToFlat, R2DToFlat, RCall, RVecLoad, RVecStore, RVecBinOp, RVecUnOp, RVecSplat,
        RVecReduce, RArrayFree
    } kind = RAtom;
    BinOp bop = BinOp::CAdd;
    UnOp uop = UnOp::CNot;
    std::vector<Atom> args;  // operands in source order
    int64_t proj = 0;
    std::string name;  // RCall function
};

struct IndexExpr {
    enum Kind { IVar, IConst, IAdd, ISub, IMul, IDiv, ITuple, IProj, IFlatToNd, INdToFlat, ICall } kind = IConst;
    std::string var;
    int64_t c = 0;
    std::vector<IndexExpr> kids;
};

enum class Strategy {
    Generic, ScatterDirect, ScatterAtomicAddInt, ScatterAtomicAddFloat, ScatterPrivatizedIntAdd,
    ScatterPrivatizedFloatAdd
};
enum class Role { Plain, Fold, Map, ReductionWrapper, Reduction, MapReduction, Iterate, SegRedOuter };

struct LoopSpec {
    std::vector<std::string> iters;
    std::vector<IndexExpr> bounds;
    enum Exec { Serial, Parallel, Workshare, Vector } exec = Serial;
    Strategy strategy = Strategy::Generic;
    bool has_red = false;
    std::string red_acc;
    IndexExpr red_init;
    bool red_mul = false;  // RMul, else RAdd
    Role role = Role::Plain;
    std::vector<IndexExpr> origins;
};

struct Stmt {
    enum Kind { SAssign, SArrayWrite, SLoop, SParallelRegion, SIf, SReturn, SBreak } kind = SBreak;
    std::string var;  // SAssign
    RHS rhs;          // SAssign
    Atom a0, a1, a2;  // SArrayWrite arr idx val; SIf cond; SReturn value
    LoopSpec loop;
    std::vector<Stmt> body, els;
};

struct Site {
    std::string name;
    std::vector<std::pair<std::string, CType>> args;
    bool has_result = false;
    std::pair<std::string, CType> result;
    std::map<std::string, CType> types;
    Stmt loop;
};

Site site_from_value(const Value& v);
std::vector<Site> read_sites(const std::string& text);

// Consecutive sites the printed C calls as one launcher (a HipGroup line:
// hgName, hgSites in program order).  The group launcher runs the members in
// one launch when they are map forms whose accesses it finds independent at
// run time, and calls them in order otherwise.
struct Group {
    std::string name;
    std::vector<std::string> members;
};
std::vector<Group> read_groups(const std::string& text);
// The group's launcher as a site (its name and the members' arguments, each
// once, in order of first appearance), for the header and signatures.
Site group_site(const Group& g, const std::vector<Site>& sites);

// ---- the GPU-schedule pass and the emitter ----------------------------------------
struct Options {
    int block = 256;          // threads per workgroup of the streaming forms
    int unroll = 8;           // elements per thread per round (map and reduction forms; even: paired)
    int waves_per_cu = 8;     // workgroups per CU of the streaming grid, x block/64
    int rows = 4;             // rows per round of the 2-D form (when its stores are deferred)
    int band_rounds = 1;      // rounds per workgroup of the 2-D form
    bool flat_maps = true;    // 1-D / N-D maps: grid sized by the data (one round per workgroup)
};

// The HIP source of every site (kernels and extern "C" launchers).
std::string emit_sites(const std::vector<Site>& sites, const std::vector<Group>& groups, const Options& opt,
                       const std::string& origin);
// The C declarations of the launchers (for the printed C to include).
std::string emit_header(const std::vector<Site>& sites);
// One line per launcher, "name result : type:arg type:arg ...", each type one
// of i64 f64 tuple array: for hosts that bind the launchers at run time.
std::string emit_signatures(const std::vector<Site>& sites);

// What the schedule pass decided for one site (also printed by --plan).
std::string describe_plan(const Site& site, const Options& opt);

}  // namespace hydgen

// main.cpp - the hydgen command line.
//
//   hydgen SITES [-o OUT.hip] [--header OUT.h] [--signatures OUT.sig] [--plan]
//          [--unroll U] [--block B] [--waves-per-cu W] [--rows R] [--band-rounds K]
//          [--map-grid flat|persistent]
//
// SITES holds one HipSite per line (the derived Show text Language.Hydrangea.
// CodegenHIP writes), and a HipGroup line per run of sites the printed C
// calls through one group launcher.  -o writes the HIP source of every site, --header the C
// declarations of their launchers, --signatures one line per launcher for
// hosts that bind them at run time, --plan prints each site's schedule.
// --unroll, --block and --waves-per-cu shape the streaming forms; --rows
// (1..8) and --band-rounds (rounds per workgroup) the 2-D form; --map-grid
// flat (the default) sizes a 1-D or N-D map's grid by the data, one round per
// workgroup, persistent caps it at --waves-per-cu per CU.
// Errors (an unsupported construct, a malformed line) print a message and
// exit with status 1: the HIP backend then stops, as the C backend stops on
// a C compiler error.
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>

#include "hydgen.hpp"

namespace {

std::string slurp(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw hydgen::Error(std::string("cannot read ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void spit(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    if (!f || !(f << text)) throw hydgen::Error("cannot write " + path);
}

int usage() {
    fprintf(stderr,
            "usage: hydgen SITES [-o OUT.hip] [--header OUT.h] [--signatures OUT.sig] [--plan] [--unroll U] [--block B] "
            "[--waves-per-cu W] [--rows R] [--band-rounds K] [--map-grid flat|persistent]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    const char* in = nullptr;
    std::string out, header, sigs;
    bool plan = false;
    hydgen::Options opt;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        if (!strcmp(a, "-o")) {
            const char* v = next();
            if (!v) return usage();
            out = v;
        } else if (!strcmp(a, "--header")) {
            const char* v = next();
            if (!v) return usage();
            header = v;
        } else if (!strcmp(a, "--signatures")) {
            const char* v = next();
            if (!v) return usage();
            sigs = v;
        } else if (!strcmp(a, "--plan")) {
            plan = true;
        } else if (!strcmp(a, "--map-grid")) {
            const char* v = next();
            if (!v || (strcmp(v, "flat") && strcmp(v, "persistent"))) return usage();
            opt.flat_maps = !strcmp(v, "flat");
        } else if (!strcmp(a, "--unroll") || !strcmp(a, "--block") || !strcmp(a, "--waves-per-cu") ||
                   !strcmp(a, "--rows") || !strcmp(a, "--band-rounds")) {
            const char* v = next();
            if (!v) return usage();
            const int x = atoi(v);
            if (!strcmp(a, "--unroll")) opt.unroll = x;
            else if (!strcmp(a, "--block")) opt.block = x;
            else if (!strcmp(a, "--rows")) opt.rows = x;
            else if (!strcmp(a, "--band-rounds")) opt.band_rounds = x;
            else opt.waves_per_cu = x;
        } else if (a[0] == '-' && a[1]) {
            return usage();
        } else {
            in = a;
        }
    }
    if (!in || opt.unroll < 1 || opt.unroll > 16 || (opt.block != 64 && opt.block != 128 && opt.block != 256 &&
                                                     opt.block != 512) || opt.waves_per_cu < 1 || opt.rows < 1 ||
        opt.rows > 8 || opt.band_rounds < 1)
        return usage();
    try {
        const std::string text = slurp(in);
        const std::vector<hydgen::Site> sites = hydgen::read_sites(text);
        const std::vector<hydgen::Group> groups = hydgen::read_groups(text);
        std::vector<hydgen::Site> callable = sites;  // what the printed C may call
        for (const auto& g : groups) callable.push_back(hydgen::group_site(g, sites));
        if (plan) {
            for (const auto& s : sites) printf("%s: %s\n", s.name.c_str(), hydgen::describe_plan(s, opt).c_str());
            for (const auto& g : groups) {
                printf("%s: group of", g.name.c_str());
                for (const auto& m : g.members) printf(" %s", m.c_str());
                printf("\n");
            }
        }
        if (!out.empty()) spit(out, hydgen::emit_sites(sites, groups, opt, in));
        if (!header.empty()) spit(header, hydgen::emit_header(callable));
        if (!sigs.empty()) spit(sigs, hydgen::emit_signatures(callable));
    } catch (const std::exception& e) {
        fprintf(stderr, "hydgen: %s\n", e.what());
        return 1;
    }
    return 0;
}

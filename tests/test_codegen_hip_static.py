"""The compiler-side modules against what they drive (CPU).  GHC is absent,
so Language.Hydrangea.CodegenHIP and HIPBackend are not compiled here; these
checks tie their text to the tested parts:

* every runtime function the offload pass emits as an RCall is declared in
  include/hyd_hip_runtime.h and exported by libhydhip.so;
* the HipSite and HipGroup records' fields are the ones hydgen reads;
* the unary operators the pass keeps on the host are the ones hydgen refuses;
* HIPBackend passes hydhip-cc the arguments tests/test_hydhip_cc.py builds
  programs with.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HS = os.path.join(ROOT, "hydrangea-compiler_amd", "haskell")
HYDGEN = os.path.join(ROOT, "hydrangea-compiler_amd", "hydgen")


def read(*p):
    with open(os.path.join(*p)) as f:
        return f.read()


CODEGEN = read(HS, "Language", "Hydrangea", "CodegenHIP.hs")
BACKEND = read(HS, "HIPBackend.hs")


def test_emitted_runtime_calls_exist():
    emitted = set(re.findall(r'"(hyd_array_\w+)"', CODEGEN))
    assert emitted == {"hyd_array_sync_host", "hyd_array_mark_host_dirty", "hyd_array_enter_host",
                       "hyd_array_alloc_device8"}
    header = read(ROOT, "include", "hyd_hip_runtime.h")
    lib = ctypes.CDLL(os.path.join(ROOT, "hydrangea-compiler_amd", "lib", "libhydhip.so"))
    for fn in emitted:
        assert re.search(r"\b%s\(" % fn, header), fn
        assert hasattr(lib, fn), fn


def test_site_fields_match_hydgen():
    fields = re.findall(r"\b(hs[A-Z]\w*) ::", CODEGEN)
    assert fields == ["hsName", "hsArgs", "hsResult", "hsTypes", "hsLoop"]
    read_fields = set(re.findall(r'field\("(hs\w+)"\)', read(HYDGEN, "cfg.cpp")))
    assert read_fields == set(fields)


def test_group_fields_match_hydgen():
    # HipGroup lines: the fields the pass shows are the ones hydgen reads,
    # and the group's call passes the members' arguments once each in order
    # of first appearance, as hydgen orders the group launcher's parameters
    fields = re.findall(r"\b(hg[A-Z]\w*) ::", CODEGEN)
    assert fields == ["hgName", "hgSites"]
    cfg = read(HYDGEN, "cfg.cpp")
    assert set(re.findall(r'field\("(hg\w+)"\)', cfg)) == set(fields) and 'is_con("HipGroup")' in cfg
    assert "args = nub (concat [as | SAssign _ (RCall _ as) <- run])" in CODEGEN
    assert "seen.emplace(a.first, a.second).second) args.push_back(a)" in read(HYDGEN, "codegen.cpp")
    assert "renderSites (hipSites arts) (hipGroups arts)" in BACKEND


def test_host_only_unary_operators_match():
    hs = set(re.search(r"op `notElem` \[([^\]]*)\]", CODEGEN).group(1).replace(" ", "").split(","))
    assert hs == {"CLog", "CSin", "CCos", "CErf"}
    cg = read(HYDGEN, "codegen.cpp")
    for op in ("CSqrt", "CExpF", "CAbsF", "CFloorF", "CCeilF", "CFloatOf", "CIntOf", "CNot", "CNeg"):
        assert "UnOp::%s" % op in cg, op
    # everything else reaches hydgen's refusal (exercised in test_hydgen_cpu.py)
    assert "have no glibc-exact device form" in cg


def test_backend_driver_arguments():
    assert 'driver = home </> "hydrangea-compiler_amd" </> "hydhip-cc"' in BACKEND
    assert '[cFile, sitesFile, "-o", exe] ++ maybe [] (\\cc -> ["--cc", cc]) mcc ++ ["--no-openmp" | not parallel]' \
        in BACKEND

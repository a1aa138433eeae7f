"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/*.h declares, its enums match the headers, its parameter defaults
match the reference's pastixInitParam, and the numerical tasks refuse to
run without a GPU (the engine has no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pastix_amd as pa
from pastix_amd import _lib, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = [os.path.join(INC, "pastix.h"), os.path.join(INC, "pastix_hip.h"),
           os.path.join(INC, "pastix", "order.h"), os.path.join(INC, "pastix", "api.h"),
           os.path.join(INC, "pastix", "solver.h"), os.path.join(INC, "pastix", "datatypes.h"),
           os.path.join(INC, "spm.h")]


def _strip_comments(txt):
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return re.sub(r"//[^\n]*", " ", txt)


def declared_functions():
    """Names of the non-inline functions declared at file scope."""
    names = set()
    for h in HEADERS:
        txt = _strip_comments(open(h).read())
        # drop preprocessor lines and inline bodies
        txt = "\n".join(l for l in txt.splitlines() if not l.lstrip().startswith("#"))
        txt = re.sub(r"static\s+inline[^{;]*\{(?:[^{}]|\{[^{}]*\})*\}", " ", txt, flags=re.S)
        # brace blocks (struct/enum bodies, extern "C" wrappers) become separators
        txt = re.sub(r"\{(?:[^{}]|\{[^{}]*\})*\}", ";", txt.replace('extern "C" {', ";"), flags=re.S)
        for stmt in txt.split(";"):
            m = re.match(r"\s*(?:extern\s+)?([A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(([^()]*)\)\s*$", stmt, flags=re.S)
            if not m:
                continue
            ret, name = m.group(1).strip(), m.group(2)
            if not ret or ret.split()[0] in ("typedef", "struct", "enum", "return"):
                continue
            names.add(name)
    return names


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    # the user API, the GPU kernel boundary and the ordering helpers
    for must in ("pastixInit", "pastix_task_numfact", "pastix_task_solve", "pastix_task_refine",
                 "gpucblk_dgemmsp", "gpucblk_sgemmsp", "gpublok_dtrsmsp", "gpublok_sgemmsp",
                 "gpucblk_dpotrfsp", "gpucblk_sgetrfsp", "pastixOrderGrid", "pastix"):
        assert must in names, must
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, "declared but not exported: %s" % missing


def _header_enums(path):
    txt = _strip_comments(open(path).read())
    vals = {}
    for body in re.findall(r"enum[^{]*\{(.*?)\}", txt, flags=re.S):
        cur = -1
        for item in body.split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                k, v = [x.strip() for x in item.split("=", 1)]
                cur = int(eval(v, {}, vals))
            else:
                k = item
                cur += 1
            vals[k] = cur
    return vals


def test_python_enums_match_the_header():
    vals = _header_enums(os.path.join(INC, "pastix", "api.h"))
    for name, v in vars(pa.iparm).items():
        if name.startswith("_"):
            continue
        assert vals["IPARM_" + name.upper()] == v, name
    for name, v in vars(pa.dparm).items():
        if name.startswith("_"):
            continue
        assert vals["DPARM_" + name.upper()] == v, name
    assert vals["IPARM_SIZE"] == _lib.IPARM_SIZE and vals["DPARM_SIZE"] == _lib.DPARM_SIZE
    assert vals["PastixFactLLT"] == pa.factotype.LLT and vals["PastixFactLU"] == pa.factotype.LU
    assert vals["PastixRefineGMRES"] == pa.refine.GMRES and vals["PastixRefineCG"] == pa.refine.CG


def test_init_param_defaults_follow_the_reference():
    """Defaults of common/api.c:420-603 (pastixInitParam) that the factor and
    solve path reads; the ordering default differs on purpose (no Scotch)."""
    ip, dp = pa.initParam()
    assert ip[pa.iparm.min_blocksize] == 160 and ip[pa.iparm.max_blocksize] == 320   # api.c:497-498
    assert ip[pa.iparm.factorization] == pa.factotype.LU                            # api.c:514
    assert ip[pa.iparm.refinement] == pa.refine.GMRES                               # api.c:537
    assert ip[pa.iparm.itermax] == 250 and ip[pa.iparm.gmres_im] == 25              # api.c:540-541
    assert ip[pa.iparm.mixed] == 0
    assert dp[pa.dparm.epsilon_refinement] == -1.                                   # api.c:585
    assert ip[pa.iparm.ordering] == pa.ordering.Personal


def test_numfact_without_gpu_fails_loudly():
    """Analysis runs on the host; the numerical factorization has no CPU
    fallback and must return an error when no HIP device is present."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    A = problems.laplacian(4, 4, 4)
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(4, 4, 4))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        assert dims[0] == 64 and len(cblk) > 0
        with pytest.raises(pa.PastixError):
            pa.task_numfact(h, spm)
    finally:
        pa.finalize(h)


def test_tuning_candidate_count_matches_the_header():
    import re
    from pastix_amd.solver import TUNE_CANDIDATES, TUNE_NCAND
    txt = open(os.path.join(ROOT, "include", "pastix.h")).read()
    m = re.search(r"#define\s+PASTIX_AMD_TUNE_NCAND\s+(\d+)", txt)
    assert m and int(m.group(1)) == TUNE_NCAND == len(TUNE_CANDIDATES)
    assert TUNE_CANDIDATES[0] == (384, 1, 1, 0)   # the default schedule is candidate 0

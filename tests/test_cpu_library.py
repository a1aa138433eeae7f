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
    assert TUNE_CANDIDATES[0] == (1 << 30, 1, 1, 0)   # the default schedule is candidate 0


def _integration_program():
    """The C program of INTEGRATION.md section 1 (its first ```c block)."""
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = txt[txt.index("## 1."):txt.index("## 2.")]
    return sec[sec.index("```c\n") + 5:sec.index("```", sec.index("```c\n") + 5)]


def test_integration_example_compiles_links_and_stops_without_gpu(tmp_path):
    """INTEGRATION.md's C caller builds with gcc against include/ and the
    library alone; without a HIP device it runs the host analysis and stops
    at the numerical factorization with PASTIX_ERR_INTERNAL (no CPU path)."""
    import subprocess
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    src = tmp_path / "app.c"
    src.write_text(_integration_program())
    exe = tmp_path / "app"
    libdir = os.path.dirname(_lib.LIBPATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", str(src), "-I", INC, "-L", libdir, "-lpastix_amd",
           "-Wl,-rpath," + libdir, "-lm", "-o", str(exe)]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        return   # the GPU run of the same program is covered by the GPU tests
    run = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert run.returncode == pa.error.INTERNAL, (run.returncode, run.stderr[-2000:])
    assert "no GPU" in run.stderr or "no HIP device" in run.stderr


def test_rhs_precision_conversions():
    """pastixRhsDoubletoSingle / pastixRhsSingleToDouble (reference
    sopalin/pastix_rhs.c:162-316): an unset handle (allocated == -1) gets a
    library-owned copy in the other precision; an entry beyond the FP32 range
    is an error, as LAPACK dlag2s reports it."""
    lib = _lib.load()
    m, n, ld = 7, 3, 9
    rng = np.random.default_rng(0)
    B = np.asfortranarray(rng.uniform(-1, 1, (ld, n)))
    dB = C.c_void_p()
    sB = C.c_void_p()
    dB2 = C.c_void_p()
    assert lib.pastixRhsInit(C.byref(dB)) == 0
    h = C.cast(dB, C.POINTER(_lib.pastix_rhs_s)).contents
    assert h.allocated == -1 and h.m == -1 and h.b is None
    h.allocated, h.flttype, h.m, h.n, h.ld, h.b = 0, 3, m, n, ld, B.ctypes.data
    assert lib.pastixRhsInit(C.byref(sB)) == 0
    assert lib.pastixRhsDoubletoSingle(dB, sB) == 0
    s = C.cast(sB, C.POINTER(_lib.pastix_rhs_s)).contents
    assert s.allocated == 1 and s.flttype == 2 and (s.m, s.n, s.ld) == (m, n, ld)
    S = np.ctypeslib.as_array(C.cast(s.b, C.POINTER(C.c_float)), shape=(ld * n,)).reshape((ld, n), order="F")
    assert np.array_equal(S[:m], B[:m].astype(np.float32))
    assert lib.pastixRhsInit(C.byref(dB2)) == 0
    assert lib.pastixRhsSingleToDouble(sB, dB2) == 0
    d = C.cast(dB2, C.POINTER(_lib.pastix_rhs_s)).contents
    assert d.allocated == 1 and d.flttype == 3
    D = np.ctypeslib.as_array(C.cast(d.b, C.POINTER(C.c_double)), shape=(ld * n,)).reshape((ld, n), order="F")
    assert np.array_equal(D[:m], B[:m].astype(np.float32).astype(np.float64))
    # wrong direction and overflow are refused
    assert lib.pastixRhsDoubletoSingle(sB, dB2) == pa.error.BADPARAMETER
    B[2, 1] = 1e300
    assert lib.pastixRhsDoubletoSingle(dB, sB) == pa.error.INTERNAL
    h.b = None
    for r in (dB, sB, dB2):
        assert lib.pastixRhsFinalize(r) == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["d", "s"])
def test_applyorder_round_trip(dtype):
    """pastix_subtask_applyorder after the (host) analysis: Forward gives
    b[peritab[i]] in place with the handle describing b, Backward restores
    the user numbering and resets the handle (reference
    sopalin/pastix_task_solve.c:169-290, bcsc/bvec_zlapmr.c:816-910)."""
    lib = _lib.load()
    grid = (5, 4, 3)
    A = problems.laplacian(*grid).astype(dtype)
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(*grid))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        n, nrhs, ld = A.shape[0], 2, A.shape[0] + 3
        rng = np.random.default_rng(1)
        b0 = np.asfortranarray(rng.uniform(-1, 1, (ld, nrhs)).astype(dtype))
        b = b0.copy(order="F")
        Bp = C.c_void_p()
        assert lib.pastixRhsInit(C.byref(Bp)) == 0
        assert lib.pastix_subtask_applyorder(h, pa.dir.Forward, n, nrhs, C.c_void_p(b.ctypes.data), ld, Bp) == 0
        hb = C.cast(Bp, C.POINTER(_lib.pastix_rhs_s)).contents
        assert hb.allocated == 0 and hb.b == b.ctypes.data and (hb.m, hb.n, hb.ld) == (n, nrhs, ld)
        assert hb.flttype == (2 if dtype == np.float32 else 3)
        assert np.array_equal(b[:n], b0[:n][invp])
        assert np.array_equal(b[n:], b0[n:])
        # a numerical subtask before the factorization is refused
        assert lib.pastix_subtask_solve(h, Bp) == pa.error.BADPARAMETER
        assert lib.pastix_subtask_applyorder(h, pa.dir.Backward, n, nrhs, C.c_void_p(b.ctypes.data), ld, Bp) == 0
        assert np.array_equal(b, b0)
        assert hb.allocated == -1 and hb.b is None
        # wrong sizes are refused before anything is touched
        assert lib.pastix_subtask_applyorder(h, pa.dir.Forward, n + 1, nrhs, C.c_void_p(b.ctypes.data), ld + 1,
                                             Bp) == pa.error.BADPARAMETER
        assert lib.pastix_subtask_applyorder(None, pa.dir.Forward, n, nrhs, C.c_void_p(b.ctypes.data), ld,
                                             Bp) == pa.error.BADPARAMETER
        assert lib.pastixRhsFinalize(Bp) == 0
    finally:
        pa.finalize(h)


def test_solve_subtasks_refuse_bad_arguments_before_any_launch():
    """pastix_subtask_solve_adv / trsm / diag / refine check the handle, the
    step order and the RHS sizes first (reference
    sopalin/pastix_task_solve.c:529-535): a NULL handle, a missing
    factorization or an RHS of the wrong size returns
    PASTIX_ERR_BADPARAMETER without touching the device."""
    lib = _lib.load()
    Bp = C.c_void_p()
    assert lib.pastixRhsInit(C.byref(Bp)) == 0
    try:
        assert lib.pastix_subtask_solve_adv(None, 111, Bp) == pa.error.BADPARAMETER
        assert lib.pastix_subtask_trsm(None, 141, 122, 111, 131, Bp) == pa.error.BADPARAMETER
        assert lib.pastix_subtask_diag(None, Bp) == pa.error.BADPARAMETER
        assert lib.pastix_subtask_refine(None, Bp, Bp) == pa.error.BADPARAMETER
        grid = (4, 4, 4)
        A = problems.laplacian(*grid)
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = pa.factotype.LLT
        ip[pa.iparm.verbose] = pa.verbose.Not
        h = pa.init(ip, dp)
        try:
            spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
            pa.task_analyze(h, spm, pa.order.grid(*grid))
            x = np.ones(64)
            hb = C.cast(Bp, C.POINTER(_lib.pastix_rhs_s)).contents
            hb.allocated, hb.flttype, hb.m, hb.n, hb.ld, hb.b = 0, 3, 64, 1, 64, x.ctypes.data
            # analysed but not factorized
            assert lib.pastix_subtask_solve_adv(h, 111, Bp) == pa.error.BADPARAMETER
            assert lib.pastix_subtask_diag(h, Bp) == pa.error.BADPARAMETER
            hb.b = None
        finally:
            pa.finalize(h)
    finally:
        lib.pastixRhsFinalize(Bp)


def test_check_param_accepts_every_refinement_method():
    lib = _lib.load()
    ip, dp = pa.initParam()
    for m in (pa.refine.GMRES, pa.refine.CG, pa.refine.SR, pa.refine.BiCGSTAB):
        ip[pa.iparm.refinement] = m
        assert lib.pastixCheckParam(C.c_void_p(ip.ctypes.data), C.c_void_p(dp.ctypes.data)) == 0
    ip[pa.iparm.refinement] = 9
    assert lib.pastixCheckParam(C.c_void_p(ip.ctypes.data), C.c_void_p(dp.ctypes.data)) == pa.error.BADPARAMETER

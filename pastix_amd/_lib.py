"""ctypes binding of ``libpastix_amd.so`` (the C-ABI declared in
``include/pastix.h`` and ``include/pastix_hip.h``).

The library is built in-tree by ``pastix_amd.build``; loading fails loudly
when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libpastix_amd.so")

pastix_int_t = C.c_int64
IPARM_SIZE = 80
DPARM_SIZE = 26


class spmatrix_t(C.Structure):
    """Mirror of ``spmatrix_t`` (include/spm.h)."""
    _fields_ = [
        ("mtxtype", C.c_int), ("flttype", C.c_int), ("fmttype", C.c_int),
        ("baseval", pastix_int_t), ("gN", pastix_int_t), ("n", pastix_int_t),
        ("gnnz", pastix_int_t), ("nnz", pastix_int_t), ("gNexp", pastix_int_t),
        ("nexp", pastix_int_t), ("gnnzexp", pastix_int_t), ("nnzexp", pastix_int_t),
        ("dof", pastix_int_t), ("dofs", C.c_void_p), ("layout", C.c_int),
        ("colptr", C.c_void_p), ("rowptr", C.c_void_p), ("loc2glob", C.c_void_p),
        ("values", C.c_void_p), ("glob2loc", C.c_void_p), ("clustnum", C.c_int),
        ("clustnbr", C.c_int), ("comm", C.c_int), ("replicated", C.c_int),
    ]


class pastix_order_t(C.Structure):
    """Mirror of ``pastix_order_t`` (include/pastix/order.h)."""
    _fields_ = [
        ("baseval", pastix_int_t), ("vertnbr", pastix_int_t), ("cblknbr", pastix_int_t),
        ("permtab", C.POINTER(pastix_int_t)), ("peritab", C.POINTER(pastix_int_t)),
        ("rangtab", C.POINTER(pastix_int_t)), ("treetab", C.POINTER(pastix_int_t)),
        ("selevtx", C.c_void_p), ("sndenbr", pastix_int_t), ("sndetab", C.c_void_p),
        ("peritab_exp", C.c_void_p),
    ]


class SolverBlok(C.Structure):
    """Mirror of ``SolverBlok`` (include/pastix/solver.h)."""
    _fields_ = [
        ("handler", C.c_void_p * 2), ("lcblknm", pastix_int_t), ("fcblknm", pastix_int_t),
        ("gbloknm", pastix_int_t), ("gfaninnm", pastix_int_t), ("frownum", pastix_int_t),
        ("lrownum", pastix_int_t), ("coefind", pastix_int_t), ("browind", pastix_int_t),
        ("inlast", C.c_int8), ("iluklvl", C.c_int), ("LRblock", C.c_void_p * 2),
    ]


class SolverCblk(C.Structure):
    """Mirror of ``SolverCblk`` (include/pastix/solver.h)."""
    _fields_ = [
        ("lock", C.c_int32), ("ctrbcnt", C.c_int32), ("cblktype", C.c_int8),
        ("partitioned", C.c_int8), ("fcolnum", pastix_int_t), ("lcolnum", pastix_int_t),
        ("fblokptr", C.POINTER(SolverBlok)), ("stride", pastix_int_t), ("lcolidx", pastix_int_t),
        ("brownum", pastix_int_t), ("brown2d", pastix_int_t), ("sndeidx", pastix_int_t),
        ("gcblknum", pastix_int_t), ("bcscnum", pastix_int_t), ("gfaninnum", pastix_int_t),
        ("lcoeftab", C.c_void_p), ("ucoeftab", C.c_void_p), ("handler", C.c_void_p * 2),
        ("selevtx", pastix_int_t), ("ownerid", C.c_int), ("threadid", C.c_int),
        ("priority", pastix_int_t),
    ]


assert C.sizeof(SolverBlok) == 104 and C.sizeof(SolverCblk) == 160


class pastix_rhs_s(C.Structure):
    """Mirror of ``struct pastix_rhs_s`` (pastix_amd/csrc/host/internal.h):
    read-only view of a right-hand-side handle made by ``pastixRhsInit``."""
    _fields_ = [
        ("allocated", C.c_int), ("flttype", C.c_int), ("m", pastix_int_t), ("n", pastix_int_t),
        ("ld", pastix_int_t), ("b", C.c_void_p),
    ]

# exported entry points: name -> (restype, argtypes)
_P = C.c_void_p
_I = pastix_int_t
_D = C.c_double
PROTOTYPES = {
    "pastixInitParam": (None, [_P, _P]),
    "pastixInit": (None, [_P, C.c_size_t, _P, _P]),
    "pastixInitWithAffinity": (None, [_P, C.c_size_t, _P, _P, _P]),
    "pastixFinalize": (None, [_P]),
    "pastix": (C.c_int, [_P, C.c_size_t, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P]),
    "pastix_task_analyze": (C.c_int, [_P, _P]),
    "pastix_task_numfact": (C.c_int, [_P, _P]),
    "pastix_task_solve": (C.c_int, [_P, _I, _I, _P, _I]),
    "pastix_task_refine": (C.c_int, [_P, _I, _I, _P, _I, _P, _I]),
    "pastix_task_solve_and_refine": (C.c_int, [_P, _I, _I, _P, _I, _P, _I]),
    "pastix_subtask_order": (C.c_int, [_P, _P, _P]),
    "pastix_subtask_symbfact": (C.c_int, [_P]),
    "pastix_subtask_reordering": (C.c_int, [_P]),
    "pastix_subtask_blend": (C.c_int, [_P]),
    "pastix_subtask_spm2bcsc": (C.c_int, [_P, _P]),
    "pastix_subtask_bcsc2ctab": (C.c_int, [_P]),
    "pastix_subtask_sopalin": (C.c_int, [_P]),
    "pastix_subtask_applyorder": (C.c_int, [_P, C.c_int, _I, _I, _P, _I, _P]),
    "pastix_subtask_trsm": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "pastix_subtask_diag": (C.c_int, [_P, _P]),
    "pastix_subtask_solve": (C.c_int, [_P, _P]),
    "pastix_subtask_refine": (C.c_int, [_P, _P, _P]),
    "pastix_subtask_solve_adv": (C.c_int, [_P, C.c_int, _P]),
    "pastixRhsInit": (C.c_int, [_P]),
    "pastixRhsFinalize": (C.c_int, [_P]),
    "pastixRhsDoubletoSingle": (C.c_int, [_P, _P]),
    "pastixRhsSingleToDouble": (C.c_int, [_P, _P]),
    "pastixCheckParam": (C.c_int, [_P, _P]),
    "pastixDumpParam": (None, [_P]),
    "pastixOrderAlloc": (C.c_int, [_P, _I, _I]),
    "pastixOrderInit": (C.c_int, [_P, _I, _I, _I, _P, _P, _P, _P]),
    "pastixOrderExit": (None, [_P]),
    "pastixOrderBase": (None, [_P, _I]),
    "pastixOrderCheck": (C.c_int, [_P]),
    "pastixOrderCopy": (C.c_int, [_P, _P]),
    "pastixOrderGet": (_P, [_P]),
    "pastixOrderGrid": (C.c_int, [_P, _I, _I, _I]),
    "pastix_amd_solve_dev": (C.c_int, [_P, C.c_int, _I, _P, _I]),
    "pastix_amd_solve_dev_noalloc": (C.c_int, [_P, C.c_int, _I, _P, _I]),
    "pastix_amd_factors_save": (C.c_int, [_P, C.c_char_p]),
    "pastix_amd_factors_load": (C.c_int, [_P, C.c_char_p]),
    "coeftab_ddump": (None, [_P, _P, C.c_char_p]),
    "coeftab_sdump": (None, [_P, _P, C.c_char_p]),
    "solverSave": (C.c_int, [_P, _P]),
    "solverLoad": (C.c_int, [_P, _P]),
    "solverExit": (None, [_P]),
    "pastix_amd_get_solver": (_P, [_P]),
    "pastix_amd_get_schedule": (C.c_int, [_P, _P, _P, _P]),
    "pastix_amd_refine_dev": (C.c_int, [_P, _I, _P, _I, _P, _I]),
    "pastix_amd_refine_dev_adv": (C.c_int, [_P, C.c_int, _I, _P, _I, _P, _I]),
    "pastix_amd_sampled_outer": (C.c_int, [_P, _P, _P, _I, _P, _I, _P, _I, C.c_double, _P]),
    "pastix_amd_values_dev": (C.c_int, [_P, _P]),
    "pastix_amd_tune": (C.c_int, [_P, _P, _P]),
    "pastix_amd_set_tuning": (C.c_int, [_P, _P]),
    "pastix_amd_spmv_dev": (C.c_int, [_P, C.c_int, _D, _P, _D, _P]),
    "pastix_amd_last_fact_seconds": (_D, [_P]),
    "pastix_amd_profile": (C.c_int, [_P, C.c_int, _P]),
    "pastix_amd_get_dims": (C.c_int, [_P, _P]),
    "pastix_amd_get_structure": (C.c_int, [_P, _P, _P, _P, _P]),
    "pastix_amd_get_coeftab": (C.c_int, [_P, C.c_int, _P]),
    # include/pastix_hip.h
    "gpucblk_dgemmsp": (_D, [C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gpucblk_sgemmsp": (_D, [C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gpublok_dgemmsp": (_D, [C.c_int, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P]),
    "gpublok_sgemmsp": (_D, [C.c_int, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P]),
    "gpublok_dtrsmsp": (_D, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _I, _P, _P, _P, _P]),
    "gpublok_strsmsp": (_D, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _I, _P, _P, _P, _P]),
    "gpucblk_dpotrfsp": (_D, [_P, _P, _D, _P, _P]),
    "gpucblk_spotrfsp": (_D, [_P, _P, _D, _P, _P]),
    "gpucblk_dgetrfsp": (_D, [_P, _P, _P, _D, _P, _P]),
    "gpucblk_sgetrfsp": (_D, [_P, _P, _P, _D, _P, _P]),
}

_lib = None


def load():
    """Load the engine; raises ``OSError`` if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise OSError("libpastix_amd.so is not built (run pastix_amd.build.build()): "
                          "the engine has no CPU fallback")
        lib = C.CDLL(LIBPATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib

"""Enumerations of the PaStiX API (same values as include/pastix/api.h,
named like the reference's Python wrapper enum.py)."""


class iparm:
    verbose = 0
    io_strategy = 1
    nnzeros = 2
    nnzeros_block_local = 3
    allocated_terms = 4
    produce_stats = 5
    trace = 6
    mc64 = 7
    ordering = 8
    ordering_default = 9
    amalgamation_lvlblas = 26
    amalgamation_lvlcblk = 27
    reordering_split = 28
    reordering_stop = 29
    splitting_strategy = 30
    min_blocksize = 36
    max_blocksize = 37
    tasks2d_level = 38
    tasks2d_width = 39
    allcand = 40
    incomplete = 41
    level_of_fill = 42
    factorization = 43
    facto_look_side = 44
    static_pivoting = 45
    free_cscuser = 46
    schur_fact_mode = 47
    transpose_solve = 48
    schur_solv_mode = 49
    applyperm_ws = 50
    refinement = 51
    nbiter = 52
    itermax = 53
    gmres_im = 54
    scheduler = 55
    thread_nbr = 56
    socket_nbr = 57
    autosplit_comm = 58
    gpu_nbr = 59
    gpu_memory_percentage = 60
    gpu_memory_block_size = 61
    global_allocation = 62
    compress_min_width = 63
    compress_min_height = 64
    compress_when = 65
    compress_method = 66
    compress_ortho = 67
    compress_reltol = 68
    compress_preselect = 69
    compress_iluk = 70
    mixed = 71
    ftz = 72
    mpi_thread_level = 73
    modify_parameter = 74
    start_task = 75
    end_task = 76
    float = 77
    mtx_type = 78
    dof_nbr = 79
    size = 80


class dparm:
    fill_in = 0
    epsilon_refinement = 1
    relative_error = 2
    epsilon_magn_ctrl = 3
    order_time = 4
    symbfact_time = 5
    reorder_time = 6
    blend_time = 7
    analyze_time = 8
    pred_fact_time = 9
    fact_time = 10
    fact_flops = 11
    fact_thflops = 12
    fact_rlflops = 13
    fact_energy = 14
    mem_fr = 15
    mem_lr = 16
    solv_time = 17
    solv_flops = 18
    solv_thflops = 19
    solv_rlflops = 20
    solv_energy = 21
    refine_time = 22
    a_norm = 23
    compress_tolerance = 24
    compress_min_ratio = 25
    size = 26


class factotype:
    POTRF = 0
    SYTRF = 1
    GETRF = 2
    PXTRF = 3
    HETRF = 4
    LLH = 0
    LDLT = 1
    LU = 2
    LLT = 3
    LDLH = 4


class ordering:
    Scotch = 0
    Metis = 1
    Personal = 2
    PtScotch = 3
    ParMetis = 4


class io:
    """IPARM_IO_STRATEGY flags (reference include/pastix/api.h)."""
    No = 0
    Load = 1
    Save = 2
    LoadGraph = 4
    SaveGraph = 8
    LoadCSC = 16
    SaveCSC = 32


class refine:
    GMRES = 0
    CG = 1
    SR = 2
    BiCGSTAB = 3


class verbose:
    Not = 0
    No = 1
    Yes = 2


class trans:
    NoTrans = 111
    Trans = 112
    ConjTrans = 113


class mtxtype:
    General = 111
    Symmetric = 112
    Hermitian = 113


class coefside:
    LCoef = 0
    UCoef = 1
    LUCoef = 2


class side:
    Left = 141
    Right = 142


class uplo:
    Upper = 121
    Lower = 122


class diag:
    NonUnit = 131
    Unit = 132


class dir:
    """pastix_dir_t (reference include/pastix/api.h)."""
    Forward = 391
    Backward = 392


class error:
    """pastix_error_t return codes (reference include/pastix/api.h:375-388)."""
    SUCCESS = 0
    UNKNOWN = 1
    ALLOC = 2
    NOTIMPLEMENTED = 3
    OUTOFMEMORY = 4
    THREAD = 5
    INTERNAL = 6
    BADPARAMETER = 7
    FILE = 8
    INTEGER_TYPE = 9
    IO = 10
    MPI = 11

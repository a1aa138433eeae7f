"""CPU checks of the analysis' static schedule (proportional mapping +
simulation, blend/propmap.c and blend/simu_run.c restated), of the oracle's
static-scheduled factorization that follows it (the CPU baseline), and of
the SolverMatrix text I/O (blend/solver_io.c).

The schedule is pinned by what makes it a schedule and by what proportional
mapping must do: every cblk runs exactly once; the lists of all threads
merge into an order in which every cblk comes after all the cblks that
update it (the waits of the threads then always end); the simulated time on
T threads lies between the one-thread time divided by T and the one-thread
time; the two subtrees below a top separator of a 2D grid go to the two
threads, one each.  The executor is pinned by the sequential
oracle's factors (themselves pinned to LAPACK / Doolittle in
test_cpu_oracle.py).  The structure I/O is pinned by a round trip through a
C caller: save, load, save again gives the same file, and the loaded
structure equals the analysed one field by field."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import _lib, problems
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _analyse(A, grid, threads, lu=False):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LU if lu else pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.thread_nbr] = threads
    h = pa.init(ip, dp)
    spm = pa.spmatrix(A, mtxtype=pa.mtxtype.General if lu else pa.mtxtype.Symmetric)
    pa.task_analyze(h, spm, pa.order.grid(*grid))
    cblk, blok, perm, invp, dims = pa.get_structure(h)
    tptr, ttsk = pa.get_schedule(h)
    pred = float(dp[pa.dparm.pred_fact_time])
    pa.finalize(h)
    return orc.Structure(cblk, blok), perm, tptr, ttsk, pred


@pytest.mark.parametrize("threads", [1, 2, 5, 16])
@pytest.mark.parametrize("grid", [(9, 8, 7), (40, 33, 1)], ids=["3d", "2d"])
def test_static_schedule_is_a_valid_schedule(grid, threads):
    A = problems.laplacian(*grid)
    S, perm, tptr, ttsk, pred = _analyse(A, grid, threads)
    assert len(tptr) == threads + 1 and tptr[0] == 0 and tptr[-1] == S.cblknbr
    assert np.array_equal(np.sort(ttsk), np.arange(S.cblknbr))
    # a merge of the lists consistent with every dependency exists: simulate
    # the threads' waits (each list in order, a cblk ready once all its
    # sources are done) and require that they all drain
    src = [[] for _ in range(S.cblknbr)]
    for b in range(len(S.blok)):
        lc, fc = S.blok[b, 2], S.blok[b, 3]
        if lc != fc:
            src[fc].append(lc)
    done = np.zeros(S.cblknbr, dtype=bool)
    head = list(tptr[:-1])
    progress = True
    while progress:
        progress = False
        for t in range(threads):
            while head[t] < tptr[t + 1] and all(done[s] for s in src[ttsk[head[t]]]):
                done[ttsk[head[t]]] = True
                head[t] += 1
                progress = True
    assert done.all(), "the thread lists deadlock"
    assert pred > 0


def test_simulated_time_on_t_threads_is_between_one_thread_over_t_and_one_thread():
    grid = (12, 12, 12)
    A = problems.laplacian(*grid)
    t = {}
    for threads in (1, 2, 4, 8):
        t[threads] = _analyse(A, grid, threads)[4]
    for threads in (2, 4, 8):
        assert t[1] / threads <= t[threads] * (1 + 1e-12)
        assert t[threads] <= t[1] * (1 + 1e-12)
    assert t[8] < t[1] / 2   # the subtrees do run concurrently


def test_one_thread_runs_the_cblks_in_a_topological_order_by_simulated_rank():
    grid = (7, 6, 5)
    S, perm, tptr, ttsk, pred = _analyse(problems.laplacian(*grid), grid, 1)
    pos = np.empty(S.cblknbr, dtype=np.int64)
    pos[ttsk] = np.arange(S.cblknbr)
    for b in range(len(S.blok)):
        lc, fc = S.blok[b, 2], S.blok[b, 3]
        assert lc == fc or pos[lc] < pos[fc]


def test_proportional_mapping_splits_the_top_subtrees():
    """2D grid: the last cblk is the top separator; the two subtrees below it
    (the halves of the dissection) hold about half the work each, so with
    two threads each half is mapped to one thread and no cblk of one half
    runs on the thread of the other."""
    grid = (31, 31, 1)
    A = problems.laplacian(*grid)
    S, perm, tptr, ttsk, pred = _analyse(A, grid, 2)
    owner = np.empty(S.cblknbr, dtype=np.int64)
    for t in range(2):
        owner[ttsk[tptr[t]:tptr[t + 1]]] = t
    father = np.full(S.cblknbr, -1)
    for k in range(S.cblknbr):
        fb, lb = S.cblkx[k, 2], S.cblkx[k + 1, 2]
        if lb > fb + 1:
            father[k] = S.blok[fb + 1, 3]
    # climb from the root through the chain of single sons (the split top
    # separator) to the first node with two sons
    sons = [[] for _ in range(S.cblknbr)]
    for k in range(S.cblknbr):
        if father[k] >= 0:
            sons[father[k]].append(k)
    node = S.cblknbr - 1
    while len(sons[node]) == 1:
        node = sons[node][0]
    assert len(sons[node]) == 2

    def subtree(r):
        out, st = [], [r]
        while st:
            v = st.pop()
            out.append(v)
            st.extend(sons[v])
        return out
    a, b = subtree(sons[node][0]), subtree(sons[node][1])
    assert len(set(owner[a])) == 1 and len(set(owner[b])) == 1
    assert owner[a[0]] != owner[b[0]]


@pytest.mark.parametrize("lu", [False, True], ids=["llt", "lu"])
@pytest.mark.parametrize("use_blas", [False, True], ids=["loops", "blas"])
def test_static_oracle_matches_the_sequential_oracle(lu, use_blas):
    grid = (12, 11, 10)
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    S, perm, tptr, ttsk, pred = _analyse(A, grid, 4, lu=lu)
    crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
    L, U, npiv = orc.factorize(S, A, perm, 2 if lu else 0, crit)
    Ls, Us, npivs, sec = orc.factorize_static(S, A, perm, crit, tptr, ttsk, use_blas=use_blas,
                                              factotype=2 if lu else 0)
    n = A.shape[0]
    Ld, Lsd = orc.expand_L(S, L, n), orc.expand_L(S, Ls, n)
    assert np.abs(Lsd - Ld).max() <= 1e-12 * np.abs(Ld).max()
    if lu:
        Ud, Usd = orc.expand_L(S, U, n), orc.expand_L(S, Us, n)
        assert np.abs(Usd - Ud).max() <= 1e-12 * np.abs(Ud).max()
    assert npivs == npiv and sec > 0


def test_static_oracle_refuses_a_list_that_misses_a_cblk():
    grid = (5, 5, 5)
    A = problems.laplacian(*grid)
    S, perm, tptr, ttsk, pred = _analyse(A, grid, 2)
    bad = ttsk.copy()
    bad[0] = bad[1]
    with pytest.raises(ValueError):
        orc.factorize_static(S, A, perm, 1e-300, tptr, bad)


SOLVER_IO_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include "pastix.h"

static long grid_n(void) { return 9 * 8 * 7; }

int main( int argc, char **argv )
{
    pastix_int_t iparm[IPARM_SIZE];
    double       dparm[DPARM_SIZE];
    pastix_data_t *pd = NULL;
    pastix_order_t ord, *op = &ord;
    spmatrix_t spm;
    SolverMatrix loaded;
    const SolverMatrix *s;
    FILE *f;
    pastix_int_t k, b, t, i, n = grid_n(), nx = 9, ny = 8, nz = 7, nnz = 0;
    pastix_int_t *colptr = malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *rows = malloc( 7 * n * sizeof( pastix_int_t ) );
    double *vals = malloc( 7 * n * sizeof( double ) );
    (void)argc;
    /* lower triangle of the 7-point Laplacian, column by column */
    for ( k = 0; k < n; k++ ) {
        pastix_int_t x = k % nx, y = ( k / nx ) % ny, z = k / ( nx * ny );
        colptr[k] = nnz;
        rows[nnz] = k; vals[nnz++] = 6.;
        if ( x + 1 < nx ) { rows[nnz] = k + 1; vals[nnz++] = -1.; }
        if ( y + 1 < ny ) { rows[nnz] = k + nx; vals[nnz++] = -1.; }
        if ( z + 1 < nz ) { rows[nnz] = k + nx * ny; vals[nnz++] = -1.; }
    }
    colptr[n] = nnz;
    memset( &spm, 0, sizeof( spm ) );
    spm.mtxtype = SpmSymmetric; spm.flttype = SpmDouble; spm.fmttype = SpmCSC;
    spm.gN = spm.n = n; spm.gnnz = spm.nnz = nnz; spm.dof = 1;
    spm.gNexp = spm.nexp = n; spm.gnnzexp = spm.nnzexp = nnz;
    spm.colptr = colptr; spm.rowptr = rows; spm.values = vals;
    pastixInitParam( iparm, dparm );
    iparm[IPARM_VERBOSE] = PastixVerboseNot;
    iparm[IPARM_THREAD_NBR] = 3;
    pastixInit( &pd, MPI_COMM_WORLD, iparm, dparm );
    if ( pd == NULL ) return 10;
    memset( &ord, 0, sizeof( ord ) );
    if ( pastixOrderGrid( &op, nx, ny, nz ) ) return 11;
    if ( pastix_subtask_order( pd, &spm, &ord ) ) return 12;
    if ( pastix_subtask_symbfact( pd ) || pastix_subtask_reordering( pd ) || pastix_subtask_blend( pd ) ) return 13;
    s = pastix_amd_get_solver( pd );
    f = fopen( argv[1], "w" );
    if ( !f || solverSave( s, f ) ) return 14;
    fclose( f );
    f = fopen( argv[1], "r" );
    if ( !f || solverLoad( &loaded, f ) ) return 15;
    fclose( f );
    f = fopen( argv[2], "w" );
    if ( !f || solverSave( &loaded, f ) ) return 16;
    fclose( f );
    /* field by field */
    if ( loaded.cblknbr != s->cblknbr || loaded.bloknbr != s->bloknbr || loaded.nodenbr != s->nodenbr ||
         loaded.coefnbr != s->coefnbr || loaded.gemmmax != s->gemmmax || loaded.thrdnbr != 3 ||
         loaded.tasknbr != s->tasknbr || loaded.brownbr != s->brownbr || loaded.colmax != s->colmax ||
         loaded.blokmax != s->blokmax || loaded.cblkmaxblk != s->cblkmaxblk )
        return 20;
    for ( k = 0; k <= s->cblknbr; k++ ) {
        const SolverCblk *a = s->cblktab + k, *c = loaded.cblktab + k;
        if ( a->fcolnum != c->fcolnum || a->lcolnum != c->lcolnum ||
             ( a->fblokptr - s->bloktab ) != ( c->fblokptr - loaded.bloktab ) || a->brownum != c->brownum )
            return 21;
        if ( k < s->cblknbr && ( a->stride != c->stride || a->ctrbcnt != c->ctrbcnt || a->cblktype != c->cblktype ) )
            return 22;
    }
    for ( b = 0; b < s->bloknbr; b++ ) {
        const SolverBlok *a = s->bloktab + b, *c = loaded.bloktab + b;
        if ( a->frownum != c->frownum || a->lrownum != c->lrownum || a->fcblknm != c->fcblknm ||
             a->lcblknm != c->lcblknm || a->coefind != c->coefind || a->browind != c->browind )
            return 23;
    }
    for ( b = 0; b < s->brownbr; b++ )
        if ( s->browtab[b] != loaded.browtab[b] ) return 24;
    for ( t = 0; t < s->tasknbr; t++ )
        if ( s->tasktab[t].prionum != loaded.tasktab[t].prionum || s->tasktab[t].cblknum != loaded.tasktab[t].cblknum ||
             s->tasktab[t].ctrbcnt != loaded.tasktab[t].ctrbcnt )
            return 25;
    for ( t = 0; t < 3; t++ ) {
        if ( s->ttsknbr[t] != loaded.ttsknbr[t] ) return 26;
        for ( i = 0; i < s->ttsknbr[t]; i++ )
            if ( s->ttsktab[t][i] != loaded.ttsktab[t][i] ) return 27;
    }
    solverExit( &loaded );
    /* a damaged file is refused */
    f = fopen( argv[3], "r" );
    if ( !f ) return 30;
    if ( solverLoad( &loaded, f ) != PASTIX_ERR_FILE ) return 31;
    fclose( f );
    pastixOrderExit( &ord );
    pastixFinalize( &pd );
    free( colptr ); free( rows ); free( vals );
    return 0;
}
"""


def test_solver_structure_text_round_trip(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    src = tmp_path / "io.c"
    src.write_text(SOLVER_IO_PROGRAM)
    exe = tmp_path / "io"
    libdir = os.path.dirname(_lib.LIBPATH)
    res = subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", str(src), "-I", INC, "-L", libdir,
                          "-lpastix_amd", "-Wl,-rpath," + libdir, "-lm", "-o", str(exe)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout
    # the damaged file: a block row range reversed
    f1, f2, bad = tmp_path / "s1.txt", tmp_path / "s2.txt", tmp_path / "bad.txt"
    run = subprocess.run([str(exe), str(f1), str(f2), str(tmp_path / "missing")], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=120)
    # the first run stops at the missing damaged file, after the round trip
    assert run.returncode == 30, (run.returncode, run.stderr[-2000:])
    assert f1.read_text() == f2.read_text()
    lines = f1.read_text().split("\n")
    assert lines[0] == "2"
    cblknbr, bloknbr, nodenbr, baseval = map(int, lines[1].split())
    assert nodenbr == 9 * 8 * 7 and baseval == 0 and bloknbr >= cblknbr
    first_blok = lines[2 + cblknbr].split()
    damaged = lines[:]
    damaged[2 + cblknbr] = "%s\t%s\t%s" % (first_blok[1], first_blok[0], first_blok[2]) \
        if first_blok[0] != first_blok[1] else "%d\t%s\t%s" % (int(first_blok[1]) + 1, first_blok[1], first_blok[2])
    bad.write_text("\n".join(damaged))
    run = subprocess.run([str(exe), str(f1), str(f2), str(bad)], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stderr[-2000:])

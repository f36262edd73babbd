"""GPU parity: libins2d_hip (through its C ABI) against the CPU oracle on the
same inputs, and against the reference-held fixtures.

Tolerances: the plain FP64 path must match within 1e-12 relative
(velocities; pressure modulo its constant for enclosed flows).  Pressure-CG
runs use a tight CG tolerance here (the oracle solves directly), so the
comparison is not limited by the solver tolerance."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from ins2d_amd import host
from ins2d_amd.cases import load
from oracle import ins2d_oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _say(*a):
    print('[gpu-test]', *a, flush=True)


def _case_periodic_noise(N=32, order4=0, Re=500.0, seed=3, **kw):
    return load('taylor_green', Nx=N, Ny=N, Re=Re, order4=order4, noise=1e-2, seed=seed, dt=0.01, t_end=0.05, **kw)


def _wall_case(kind, Nx, Ny, **kw):
    c = load('LDC_unsteady', Nx=Nx, Ny=Ny, **kw)
    if kind == 'outflow':
        c.BCtype = lambda: {'BC_unsteady': 0, 'u': dict(left='dir', right='pres', low='dir', up='dir'),
                            'v': dict(left='dir', right='sym', low='dir', up='dir')}
        c.uBC = lambda x, y, t, o: np.ravel(np.add.outer(0 * np.atleast_1d(x), 4 * np.atleast_1d(y) * (1 - np.atleast_1d(y))) *
                                            (np.abs(np.atleast_1d(x)) < 1e-12)[:, None], order='F')
    elif kind == 'channel':
        c.BCtype = lambda: {'BC_unsteady': 0, 'u': dict(left='per', right='per', low='dir', up='dir'),
                            'v': dict(left='per', right='per', low='dir', up='dir')}
    return c


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('name,kernels', [('ldc', 0), ('outflow', 0), ('periodic2', 0), ('periodic2', 1), ('periodic4', 0)])
def test_rhs_matches_oracle(name, kernels):
    _say('rhs', name, kernels)
    if name == 'ldc':
        case = _wall_case('ldc', 24, 20, poisson=3, CG_acc=1e-13, kernels=kernels)
    elif name == 'outflow':
        case = _wall_case('outflow', 30, 12, poisson=3, CG_acc=1e-13, sx=1.04, kernels=kernels)
    else:
        case = _case_periodic_noise(32, order4=1 if name == 'periodic4' else 0, kernels=kernels)
    opts = O.build(case)
    solver, hopts, V0, p0 = host.setup(case)
    rng = np.random.default_rng(7)
    V = rng.standard_normal(opts.grid.NV)
    p = rng.standard_normal(opts.grid.Np)
    mr, F = solver.rhs(V, p)
    mro, Fo = O.F(V, V, p, 0.0, opts)
    assert _rel(F, Fo) < RTOL
    assert abs(mr - mro) <= RTOL * mro
    solver.close()


@pytest.mark.parametrize('Nx,Ny', [(1024, 1024), (4096, 128), (128, 4096), (2048, 32), (16, 16), (512, 256),
                                   (64, 48), (48, 64)])
def test_fft_solve_register_transform_lengths(Nx, Ny):
    """The power-of-two FFT kernels pick 4, 8 or 16 elements per thread by
    length and kernel (row transforms: 16 at 4096, 4 at 1024, else 8; column
    solve: 4 at 1024 and 4096, else 8); each choice and each remainder pass
    (log2 N mod 2, 3, 4) against the oracle's numpy FFT solve; (64, 48) and
    (48, 64) mix them with the LDS transforms and explicit transposes of the
    other lengths (row-major spectra).  The solve's
    condition number grows like N^2, so the large grids are compared at
    1e-11 (max norm, relative) and the residual at 1e-10 max|f|."""
    _say('fft lengths', Nx, Ny)
    case = load('taylor_green', Nx=Nx, Ny=Ny, Re=500.0, noise=1e-2, seed=3)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(opts.grid.Np)
    f -= f.mean()
    dp, _, _ = solver.poisson(f)
    dpo = O.pressure_poisson(f, 0.0, opts)
    r = opts.discretization.A @ dp - f
    _say('fft lengths', Nx, Ny, 'rel', _rel(dp, dpo), 'residual', np.abs(r).max() / np.abs(f).max())
    assert _rel(dp, dpo) < (1e-12 if Nx * Ny <= 1 << 17 else 1e-11)
    assert np.abs(r).max() < 1e-10 * np.abs(f).max()
    solver.close()


@pytest.mark.parametrize('name', ['fft2', 'fft4_ref', 'fft4_exact', 'fft_nonpow2', 'cg_ldc', 'cg_outflow', 'cg_large',
                                  'direct_ldc', 'direct_outflow', 'direct_channel', 'direct_bfs'])
def test_poisson_solves_assembled_system(name):
    _say('poisson', name)
    fft_symbol = 0
    if name == 'fft2':
        case = _case_periodic_noise(64)
    elif name.startswith('fft4'):
        case = _case_periodic_noise(64, order4=1, fft_symbol=1 if name == 'fft4_exact' else 0)
        fft_symbol = 1 if name == 'fft4_exact' else 0
    elif name == 'fft_nonpow2':
        case = load('shear_layer', Nx=40, Ny=24)
    elif name == 'cg_ldc':
        case = _wall_case('ldc', 40, 30, poisson=3, CG_acc=1e-12)
    elif name == 'cg_outflow':
        case = _wall_case('outflow', 48, 20, poisson=3, CG_acc=1e-12, sx=1.03)
    elif name == 'cg_large':
        case = _wall_case('ldc', 256, 160, poisson=3, CG_acc=1e-11, CG_maxit=5000)
    elif name == 'direct_ldc':
        case = _wall_case('ldc', 40, 30, poisson=1, sx=1.05, sy=1.05)
    elif name == 'direct_outflow':
        case = _wall_case('outflow', 48, 20, poisson=1, sx=1.03)
    elif name == 'direct_channel':
        case = _wall_case('channel', 36, 20, poisson=1, sy=1.04)
    else:
        case = load('BFS_unsteady', Nx=256, Ny=64, poisson=1)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(opts.grid.Np)
    singular = name not in ('cg_outflow', 'direct_outflow', 'direct_bfs')
    if singular:
        f -= f.mean()
    dp, it, rel = solver.poisson(f)
    A = opts.discretization.A
    r = A @ dp - f
    if name == 'fft4_ref':
        # the reference's FFT branch always uses the second-order symbol
        # (operator_divergence.m:283), so the fourth-order system is not
        # solved exactly; compare with the oracle's own FFT branch instead
        dpo = O.pressure_poisson(f, 0.0, opts)
        assert _rel(dp, dpo) < 1e-12
    elif name.startswith('fft'):
        assert np.abs(r).max() < 1e-11 * np.abs(f).max()
        # the oracle's exact symbol is the DFT of its assembled A
        dpo = O.pressure_poisson(f, 0.0, opts)
        assert _rel(dp, dpo) < 1e-12
    elif name.startswith('direct'):
        # fast diagonalization against the oracle's sparse LU (pressure
        # modulo its constant where A is singular); normwise backward error
        # at the level of a backward-stable direct solve
        An = abs(A).sum(axis=1).max()
        berr = np.abs(r).max() / (An * np.abs(dp).max() + np.abs(f).max())
        _say('direct', name, 'backward error', berr)
        assert berr < 1e-14
        assert np.abs(r).max() < 1e-11 * np.abs(f).max()
        dpo = O.pressure_poisson(f, 0.0, opts)
        if singular:
            dp, dpo = dp - dp.mean(), dpo - dpo.mean()
        _say('direct', name, 'rel', _rel(dp, dpo))
        assert _rel(dp, dpo) < 1e-11
    else:
        tol = {'cg_ldc': 1e-12, 'cg_outflow': 1e-12, 'cg_large': 1e-11}[name]
        assert rel <= tol and it > 0
        assert np.linalg.norm(r) <= 2 * tol * np.linalg.norm(f)
    solver.close()


def test_cg_nonconvergence_is_reported():
    _say('cg nonconvergence')
    case = _wall_case('ldc', 32, 32, poisson=3, CG_acc=1e-14, CG_maxit=3)
    solver, _, _, _ = host.setup(case)
    f = np.random.default_rng(1).standard_normal(solver.Np)
    f -= f.mean()
    with pytest.raises(host.Ins2dError) as e:
        solver.poisson(f)
    assert e.value.code == 3
    solver.close()


def _compare_runs(case, nsteps, const_mode_p):
    opts = O.build(case)
    Vo, po, rows_o, _ = O.run(case, nsteps=nsteps)
    V, p, rows = host.main(case, nsteps=nsteps)
    assert _rel(V, Vo) < RTOL
    if const_mode_p:
        p = p - p.mean()
        po = po - po.mean()
    _say('run parity', f'V {_rel(V, Vo):.2e} p {_rel(p, po):.2e}')
    assert _rel(p, po) < RTOL
    ro, rg = np.asarray(rows_o), np.asarray(rows)
    for c in (3, 7):   # maxres, kinetic energy
        assert np.all(np.abs(rg[:, c] - ro[:, c]) <= RTOL * np.abs(ro[:, c]).max())
    return rows


@pytest.mark.parametrize('poisson', [1, 3])
def test_ldc_run_matches_oracle_and_reference_profiles(poisson):
    """results/LDC_unsteady_1.000e+02_20x20_0: RK33P2, Re=100, 100 steps,
    with the reference's direct solve (fast diagonalization here) and with
    tight PCG."""
    _say('ldc 20x20 run', poisson)
    case = load('LDC_unsteady', RK='RK33P2', poisson=poisson, CG_acc=1e-13, CG_maxit=2000)
    rows = _compare_runs(case, 100, True)
    d = os.path.join(GOLDEN, 'LDC_unsteady_1.000e+02_20x20_0')
    ref = np.loadtxt(os.path.join(d, 'convergence.txt'), skiprows=3)
    rows = np.asarray(rows)
    for c in (3, 7):
        sel = np.abs(ref[:, c]) > 0
        assert np.max(np.abs(rows[sel, c] / ref[sel, c] - 1)) < 1.2e-8


@pytest.mark.parametrize('kernels', [0, 1])
def test_shear_layer_run_matches_oracle_and_reference_log(kernels):
    """results/shear_layer01: inviscid, 40x40 (FFT of a non-power-of-two
    length), RK44, u shifted by 1; first 50 steps."""
    _say('shear layer 40x40 run', kernels)
    case = load('shear_layer', Nx=40, Ny=40, u_shift=1.0, kernels=kernels)
    rows = _compare_runs(case, 50, False)
    ref = np.loadtxt(os.path.join(GOLDEN, 'shear_layer01', 'convergence.txt'), skiprows=3)[:51]
    rows = np.asarray(rows)
    for c in (3, 5, 7):
        assert np.max(np.abs(rows[:, c] / ref[:, c] - 1)) < 1.2e-8


@pytest.mark.parametrize('poisson', [1, 3])
def test_outflow_stretched_run_matches_oracle(poisson):
    _say('outflow stretched run', poisson)
    case = _wall_case('outflow', 32, 16, poisson=poisson, CG_acc=1e-13, CG_maxit=4000, sx=1.03, Re=200.0, dt=0.005)
    _compare_runs(case, 20, False)


def test_bfs_run_matches_oracle():
    """Backward-facing step (BFS_unsteady case files: step inflow, 'pres'
    outflow), tanh-stretched 128x32, the reference's direct solve."""
    _say('bfs 128x32 run')
    case = load('BFS_unsteady', Nx=128, Ny=32, poisson=1)
    _compare_runs(case, 10, False)


@pytest.mark.parametrize('order4,rk', [(0, 'RK44'), (1, 'RK44'), (0, 'RK44C23'), (0, 'DOPRI6')])
def test_periodic_run_matches_oracle(order4, rk):
    _say('periodic run', order4, rk)
    # fourth order with the exact FFT symbol (the reference's second-order
    # symbol leaves the fourth-order divergence and the run blows up)
    case = _case_periodic_noise(64, order4=order4, RK=rk, fft_symbol=order4)
    _compare_runs(case, 10, False)


@pytest.mark.parametrize('Nx,Ny,kernels', [(128, 32, 0), (32, 128, 0), (64, 32, 1)])
def test_periodic_rectangular_run_matches_oracle(Nx, Ny, kernels):
    """Periodic runs with hx != hy through the fused ends of the FFT solve:
    on the second-order periodic kernels (kernels=0) the first row transform
    computes the divergence of the stage velocity itself, and on every path
    the additional solve's last transform adds dp into p."""
    _say('periodic rectangular run', Nx, Ny, kernels)
    case = load('taylor_green', Nx=Nx, Ny=Ny, Re=500.0, noise=1e-2, seed=3, dt=0.01, t_end=0.05, kernels=kernels)
    _compare_runs(case, 10, False)


@pytest.mark.parametrize('Nx,Ny', [(128, 48), (256, 40), (96, 64)])
def test_fourth_order_pair_kernels_match_oracle(Nx, Ny):
    """Fourth-order periodic runs where Nx is a multiple of 128 (the RHS takes
    two cells per thread over 128-wide strips, the strips wrapping around in
    x and shorter than their 64 rows in y) and where it is not (96: one cell
    per thread); the projection takes two cells per thread for every even
    Nx."""
    _say('fourth-order pair kernels', Nx, Ny)
    case = load('taylor_green', Nx=Nx, Ny=Ny, Re=500.0, noise=1e-2, seed=5, dt=0.005, t_end=0.05, order4=1,
                fft_symbol=1)
    _compare_runs(case, 8, False)


def test_large_periodic_invariants():
    """1024^2 Taylor-Green with noise, inviscid: after steps the field stays
    discretely divergence-free, total momentum is conserved to round-off and
    the kinetic energy drift over 20 steps is tiny (RK4 dissipates only at
    O(dt^4) per unit time)."""
    _say('1024^2 invariants')
    case = load('taylor_green', Nx=1024, Ny=1024, Re=1e100)
    solver, opts, V0, p0 = host.setup(case)
    solver.set_state(V0, p0)
    solver.input_check(0.0, 1)
    d0 = solver.diagnostics(0.0)
    assert d0[1] < 1e-12
    solver.step(0.0, opts.dt, 20)
    d1 = solver.diagnostics(20 * opts.dt)
    assert d1[1] < 1e-12
    assert abs(d1[2] - d0[2]) < 1e-10 and abs(d1[3] - d0[3]) < 1e-10
    assert abs(d1[4] - d0[4]) / d0[4] < 1e-9
    solver.close()


def test_fp64_matrix_core_result_layout():
    """The direct solve's GEMM reads the accumulator of lane l, register r as
    D[l / 16 + 4 r][l % 16]; integer operands make every product exact."""
    import ctypes
    from ins2d_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(0)
    A = rng.integers(-9, 10, (16, 4)).astype(np.float64)
    B = rng.integers(-9, 10, (4, 16)).astype(np.float64)
    D = np.zeros(256)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.ins2d_fdm_mfma_probe(P(np.ascontiguousarray(A)), P(np.ascontiguousarray(B)), P(D)) == 0
    ref = A @ B
    raw = D.reshape(64, 4)
    mapped = np.array([[raw[l, r] for r in range(4)] for l in range(64)])
    expect = np.array([[ref[(l // 16) + 4 * r, l % 16] for r in range(4)] for l in range(64)])
    _say('fp64 mfma layout as assumed:', np.array_equal(mapped, expect))
    assert np.array_equal(mapped, expect)


@pytest.mark.parametrize('name', ['ldc', 'outflow', 'channel', 'bfs'])
def test_ic0_pcg_matches_cg_matlab(name):
    """IC(0)-PCG on the GPU (wavefront triangular solves) against the oracle's
    cg_matlab: the same iteration count (within one) and the same solution."""
    _say('ic0 pcg', name)
    kw = dict(poisson=4, CG_acc=1e-12, CG_maxit=5000, CG_precond='ic0')
    if name == 'ldc':
        case = _wall_case('ldc', 40, 30, sx=1.03, **kw)
    elif name == 'outflow':
        case = _wall_case('outflow', 48, 20, sx=1.03, **kw)
    elif name == 'channel':
        case = _wall_case('channel', 36, 20, sy=1.04, **kw)
    else:
        case = load('BFS_unsteady', Nx=128, Ny=32, **kw)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(8)
    f = rng.standard_normal(opts.grid.Np)
    if name in ('ldc', 'channel'):
        f -= f.mean()
    dp, it, rel = solver.poisson(f)
    xo, ito, _ = O.cg_matlab(opts.discretization.A, f, 1e-12, np.zeros(opts.grid.Np), 5000)
    _say(name, 'iterations gpu', it, 'oracle', ito, 'rel diff', _rel(dp, xo))
    assert abs(it - ito) <= 1
    assert _rel(dp, xo) < 1e-9
    solver.close()


@pytest.mark.parametrize('name', ['ldc', 'outflow', 'channel', 'bfs'])
def test_fdm_pcg_meets_the_plain_stopping_rule(name):
    """PCG with the fast-diagonalization inverse as preconditioner
    (cg_precond 3): the stopping rule of cg.F:247 at 1e-10 within two
    iterations, the true residual within 1.5x of it, and the solution of the
    oracle's sparse LU (modulo the constant where A is singular)."""
    _say('fdm pcg', name)
    def make(**kw):
        if name == 'ldc':
            return _wall_case('ldc', 40, 30, sx=1.03, **kw)
        if name == 'outflow':
            return _wall_case('outflow', 48, 20, sx=1.03, **kw)
        if name == 'channel':
            return _wall_case('channel', 36, 20, sy=1.04, **kw)
        return load('BFS_unsteady', Nx=256, Ny=64, **kw)
    case = make(poisson=3, CG_acc=1e-10, CG_maxit=5000, CG_precond='fdm')
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(opts.grid.Np)
    singular = name in ('ldc', 'channel')
    if singular:
        f -= f.mean()
    dp, it, rel = solver.poisson(f)
    r_true = np.linalg.norm(opts.discretization.A @ dp - f) / np.linalg.norm(f)
    dpo = O.pressure_poisson(f, 0.0, O.build(make(poisson=1)))
    if singular:
        dp, dpo = dp - dp.mean(), dpo - dpo.mean()
    _say(name, 'iterations', it, 'relres', rel, 'true', r_true, 'vs LU', _rel(dp, dpo))
    assert 1 <= it <= 2 and rel <= 1e-10 and r_true <= 1.5e-10
    assert _rel(dp, dpo) < 1e-9
    solver.close()


def test_fdm_pcg_second_iteration():
    """At 1e-14 the fast-diagonalization PCG on the stretched 40x30 cavity
    takes a second iteration (the preconditioner apply and direction update
    after the first convergence test run) and reaches the tolerance, with
    the solution of the oracle's sparse LU."""
    _say('fdm pcg second iteration')
    case = _wall_case('ldc', 40, 30, sx=1.03, poisson=3, CG_acc=1e-14, CG_maxit=30, CG_precond='fdm')
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    ref = O.build(_wall_case('ldc', 40, 30, sx=1.03, poisson=1))
    for seed in range(3):
        f = np.random.default_rng(seed).standard_normal(opts.grid.Np)
        f -= f.mean()
        dp, it, rel = solver.poisson(f)
        dpo = O.pressure_poisson(f, 0.0, ref)
        _say('seed', seed, 'iterations', it, 'relres', rel, 'vs LU', _rel(dp - dp.mean(), dpo - dpo.mean()))
        assert it == 2 and rel <= 1e-14
        assert _rel(dp - dp.mean(), dpo - dpo.mean()) < 1e-9
    solver.close()


@pytest.mark.parametrize('acc,nsteps', [(1e-10, 7), (1e-14, 5)])
def test_fdm_pcg_device_iteration_mode(acc, nsteps):
    """ins2d_step with the fast-diagonalization PCG runs one iteration per
    solve and checks the stopping rule once per call on the device; with
    profiling on, every solve is checked on the host.  Called one step at a
    time (several steps per call also share the F between steps, which
    rounds differently: test_steps_in_one_call_match_single_steps), both
    give the same V and p bit for bit and the same iteration total: at 1e-10
    every solve meets the rule after one iteration (no restart), at 1e-14
    the first solve needs a second one and the call starts over from its
    saved state on the host-checked path."""
    _say('fdm pcg device iteration', acc, nsteps)
    case = _wall_case('ldc', 40, 30, sx=1.03, poisson=3, CG_acc=acc, CG_maxit=30, CG_precond='fdm')
    out = []
    for prof in (True, False):
        solver, opts, V0, p0 = host.setup(case)
        solver.set_state(V0, p0)
        solver.input_check(opts.t_start)
        solver.profile(prof)
        dt = opts.dt
        total0 = solver.solver_stats()[2]   # the initial pressure solve
        for n in range(nsteps + 2):
            solver.step(opts.t_start + n * dt, dt, 1)
        V, p = solver.get_state()
        it, rel, total = solver.solver_stats()
        _say('profile', prof, 'last it', it, 'relres', rel, 'total', total - total0)
        out.append((V, p, it, total - total0))
        solver.close()
    (Va, pa, ita, ta), (Vb, pb, itb, tb) = out
    assert np.array_equal(Va, Vb) and np.array_equal(pa, pb)
    assert ta == tb and ita == itb
    solves = (nsteps + 2) * 4   # RK44: three stage solves and the additional solve per step
    assert ta == solves if acc == 1e-10 else ta > solves


def test_ldc_run_fdm_pcg_matches_oracle():
    """LDC 20x20 fixture configuration, PCG with the fast-diagonalization
    preconditioner on the GPU against the oracle's direct solve."""
    _say('ldc 20x20 run fdm pcg')
    case = load('LDC_unsteady', RK='RK33P2', poisson=3, CG_acc=1e-13, CG_maxit=2000, CG_precond='fdm')
    _compare_runs(case, 30, True)


def test_ldc_convergence_file_and_restart(tmp_path):
    """host.main writes convergence.txt in the reference's layout: against
    the LDC 20x20 fixture (the test_outputs comparison), and a run resumed
    from a restart file (restart_n, restart_from) ends bit-identical to the
    run that went straight through."""
    from test_outputs import compare_with_fixture
    _say('ldc 20x20 convergence.txt and restart')
    case = load('LDC_unsteady', RK='RK33P2', poisson=1)
    d = str(tmp_path / 'run')
    host.main(case, nsteps=100, path_results=d)
    compare_with_fixture(os.path.join(d, 'convergence.txt'), slack=True)
    d2 = str(tmp_path / 'restart')
    V1, p1, _ = host.main(case, nsteps=9, path_results=d2, restart_n=5, snapshots=True)
    assert os.path.exists(os.path.join(d2, 'restart_00000005.mat'))
    assert os.path.exists(os.path.join(d2, 'snapshots', 'Vp_n=10.mat'))
    V2, p2, rows = host.main(case, nsteps=5, restart_from=os.path.join(d2, 'restart_00000005.mat'))
    assert rows[-1][0] == 10
    assert np.array_equal(V1, V2) and np.array_equal(p1, p2)


def test_ldc_run_ic0_matches_oracle():
    """LDC 20x20 fixture configuration with the reference's cg_matlab
    preconditioner (poisson = 4, IC(0)) on both sides."""
    _say('ldc 20x20 run ic0')
    case = load('LDC_unsteady', RK='RK33P2', poisson=4, CG_acc=1e-13, CG_maxit=2000, CG_precond='ic0')
    _compare_runs(case, 30, True)


@pytest.mark.parametrize('name', ['ldc', 'bfs'])
def test_pressure_rhs_and_residual_match_oracle(name):
    """ins2d_pressure_rhs: f = M Om^-1 F(V, 0) (pressure_additional_solve.m:19-22
    with p = 0, ydM = 0 for steady boundary conditions) on the resident
    velocity; ins2d_residual: ||f - A x|| / ||f||."""
    _say('pressure rhs', name)
    case = (_wall_case('ldc', 24, 20, sx=1.04, poisson=1) if name == 'ldc'
            else load('BFS_unsteady', Nx=64, Ny=32, poisson=1))
    opts = O.build(case)
    solver, _, V0, p0 = host.setup(case)
    rng = np.random.default_rng(12)
    V = V0 + 0.1 * rng.standard_normal(V0.shape)
    solver.set_state(V, p0)
    f = solver.pressure_rhs(0.0)
    g, d = opts.grid, opts.discretization
    _, R = O.F(V, V, np.zeros(g.Np), 0.0, opts)
    fo = d.M @ (g.Om_inv * R) + d.ydM
    assert _rel(f, fo) < 1e-12
    x = rng.standard_normal(g.Np)
    ro = np.linalg.norm(f - d.A @ x) / np.linalg.norm(f)
    assert abs(solver.residual(f, x) - ro) <= 1e-12 * ro
    solver.close()


@pytest.mark.parametrize('RK,poisson,p_add', [('RK44', 1, 1), ('RK33P2', 3, 1), ('RK44', 1, 0)])
def test_tg_temporal_unsteady_bc_matches_oracle(RK, poisson, p_add):
    """case_files/TG_temporal: Dirichlet values that change in time
    (BC_unsteady = 1), 20x20, dt = 0.01 to t = 1 through host.main (one
    ins2d_step_unsteady call per step): the oracle's run step by step, and
    with RK44 the reference-held solution at t = 1 (dt = 1e-3, within the
    RK44 time error of dt = 0.01).  p_add = 0 takes the last stage pressure
    (time_ERK.m:162)."""
    from scipy.io import loadmat
    _say('tg_temporal unsteady bc', RK, poisson, p_add)
    case = load('TG_temporal', RK=RK, poisson=poisson, dt=0.01, CG_acc=1e-13, CG_maxit=2000, p_add_solve=p_add)
    _compare_runs(case, 100, True)
    if RK == 'RK44' and p_add:
        V, p, _ = host.main(case, diagnostics=False)
        m = loadmat(os.path.join(GOLDEN, 'TG_temporal', 'TG_Re100_N20_dt1e-3_Dir.mat'))
        pf = m['p'].ravel()
        assert np.max(np.abs(V - m['V'].ravel())) < 1e-9
        assert np.max(np.abs((p - p.mean()) - (pf - pf.mean()))) < 1e-9


def test_unsteady_step_chunks_match_single_steps():
    """One ins2d_step_unsteady call over 150 steps (three staging chunks of
    at most 64 steps) ends bit-identical to 150 calls of one step, and leaves
    the handle's boundary state at the final time (diagnostics agree)."""
    _say('tg_temporal unsteady chunks')
    case = load('TG_temporal', dt=0.005, poisson=1)
    out = []
    for chunked in (False, True):
        solver, opts, V0, p0 = host.setup(case)
        solver.set_state(V0, p0)
        solver.input_check(0.0, 1)
        t, dt = 0.0, opts.dt
        if chunked:
            solver.step_unsteady(t, dt, 150, *host.unsteady_samples(case, opts, solver._c, t, dt, 150))
            for _ in range(150):
                t = t + dt
        else:
            for _ in range(150):
                solver.step_unsteady(t, dt, 1, *host.unsteady_samples(case, opts, solver._c, t, dt, 1))
                t = t + dt
        V, p = solver.get_state()
        out.append((V, p, solver.diagnostics(t)))
        solver.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_unsteady_step_rejects_periodic_kernels():
    _say('unsteady step on periodic kernels')
    solver, opts, V0, p0 = host.setup(_case_periodic_noise(16))
    with pytest.raises(host.Ins2dError) as e:
        solver.step_unsteady(0.0, 0.01, 1, np.zeros(1000), np.zeros(100))
    assert e.value.code == 4
    solver.close()


def test_direct_early_exit_reads_every_column_tile():
    """The direct solve's early exit rides on its GEMM launches: the first
    records max |f| per 32-column tile, the last zeroes dp when all of them
    are below 1e-14.  On a grid with three column tiles, f below the threshold
    everywhere but one value just above it in the last tile must be solved
    (and match the oracle); below it everywhere, dp = 0 exactly."""
    _say('direct early exit, three column tiles')
    case = _wall_case('outflow', 40, 70, poisson=1)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(19)
    f = rng.standard_normal(opts.grid.Np)
    f *= 0.5e-14 / np.abs(f).max()
    dp, _, _ = solver.poisson(f)
    assert np.all(dp == 0.0) and np.all(O.pressure_poisson(f, 0.0, opts) == 0.0)
    g = f.copy()
    g[5 + 40 * 69] = 2e-14   # column 69, the third tile
    dp, _, _ = solver.poisson(g)
    dpo = O.pressure_poisson(g, 0.0, opts)
    assert np.abs(dp).max() > 0.0
    assert _rel(dp, dpo) < 1e-12
    solver.close()


@pytest.mark.parametrize('name', ['fft_reg', 'fft_lds', 'direct', 'cg_jacobi', 'cg_ic0', 'cg_fdm'])
def test_poisson_early_exit(name):
    """pressure_poisson.m:21-39: a negligible right-hand side returns dp = 0
    exactly (FFT branch: max |fft2(f)| < 1e-14; matrix solvers: max |f| <
    1e-14), as the oracle does; just above the threshold the solve runs."""
    _say('poisson early exit', name)
    if name == 'fft_reg':
        case = _case_periodic_noise(32)
    elif name == 'fft_lds':
        case = load('shear_layer', Nx=40, Ny=24)
    elif name == 'direct':
        case = _wall_case('ldc', 24, 20, poisson=1)
    else:
        case = _wall_case('outflow', 24, 20, poisson=3, CG_acc=1e-10, CG_maxit=5000,
                          CG_precond={'cg_jacobi': 'jacobi', 'cg_ic0': 'ic0', 'cg_fdm': 'fdm'}[name])
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(opts.grid.Np)
    f -= f.mean()
    # below the threshold: |fft2(f)| <= sum |f|, so max |f| * Np < 1e-14
    tiny = f * (0.5e-14 / (np.abs(f).max() * opts.grid.Np))
    dp, _, _ = solver.poisson(tiny)
    dpo = O.pressure_poisson(tiny, 0.0, opts)
    assert np.all(dpo == 0.0) and np.all(dp == 0.0)
    # above it the solve runs (and matches the oracle); the CG modes also
    # return x0 = 0 while ||f||^2 <= tol^2 (cg.F:225-231, cg_matlab.m:61-66),
    # so the right-hand side is taken well above that too
    big = f * (1e-6 / np.abs(f).max())
    dp, _, _ = solver.poisson(big)
    dpo = O.pressure_poisson(big, 0.0, opts)
    assert np.abs(dp).max() > 0.0
    if name != 'direct' and name.startswith('fft'):
        assert _rel(dp, dpo) < 1e-11
    solver.close()


@pytest.mark.parametrize('name', ['ldc', 'outflow', 'channel'])
def test_ilu_pcg_matches_oracle(name):
    """poisson = 2 (pressure_poisson.m:52-64, pcg with the ILU(0) factor L as
    M = L L^T): the GPU's wavefront triangular solves with the IC(0) pivots
    (equal to ILU(0)'s for these symmetric 5-point matrices) reach the
    oracle's iteration count within one and its solution."""
    _say('ilu pcg', name)
    kw = dict(poisson=2, CG_acc=1e-12, CG_maxit=5000)
    if name == 'ldc':
        case = _wall_case('ldc', 40, 30, sx=1.03, **kw)
    elif name == 'outflow':
        case = _wall_case('outflow', 48, 20, sx=1.03, **kw)
    else:
        case = _wall_case('channel', 36, 20, sy=1.04, **kw)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(8)
    f = rng.standard_normal(opts.grid.Np)
    if name in ('ldc', 'channel'):
        f -= f.mean()
    dp, it, rel = solver.poisson(f)
    xo, ito = O.pcg_ilu(opts.discretization.A, f, 1e-12, 5000, opts.solversettings.L)
    _say(name, 'iterations gpu', it, 'oracle', ito, 'rel diff', _rel(dp, xo))
    assert abs(it - ito) <= 1
    assert _rel(dp, xo) < 1e-9
    solver.close()


def test_ldc_run_ilu_pcg_matches_oracle():
    """results/LDC_unsteady_1.000e+02_20x20_0 configuration with poisson = 2
    at a tight tolerance: the whole run against the oracle's."""
    _say('ldc 20x20 run ilu pcg')
    case = load('LDC_unsteady', RK='RK33P2', poisson=2, CG_acc=1e-13, CG_maxit=2000)
    _compare_runs(case, 20, True)


def _jacobi_pcg(A, f, tol, maxit):
    """Jacobi-preconditioned CG with the stopping rule of cg.F:247
    (||r||^2 / ||f||^2 <= tol^2, x0 = 0), the iteration the multi-block GPU
    solver runs: returns x and the iteration count."""
    A = A.tocsr()
    dinv = 1.0 / A.diagonal()
    x = np.zeros_like(f)
    r = f.copy()
    z = dinv * r
    p = z.copy()
    nf = f @ f
    rz = r @ z
    for it in range(1, maxit + 1):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if (r @ r) / nf <= tol * tol:
            return x, it
        z = dinv * r
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    raise RuntimeError('no convergence')


@pytest.mark.parametrize('shape', [(256, 160), (512, 256)])
def test_two_launch_pcg_matches_jacobi_pcg(shape):
    """The multi-block Jacobi-PCG (more than 32768 cells: two launches per
    iteration, alpha and beta formed by the last block of each launch)
    against the same iteration in numpy: the iteration count within one, the
    solutions within the tolerance."""
    _say('two-launch pcg', shape)
    tol = 1e-10
    case = _wall_case('ldc', shape[0], shape[1], poisson=3, CG_acc=tol, CG_maxit=20000)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(opts.grid.Np)
    f -= f.mean()
    xr, itr = _jacobi_pcg(opts.discretization.A, f, tol, 20000)
    for rep in range(2):   # the second solve replays the captured chunks
        dp, it, rel = solver.poisson(f)
        _say('two-launch pcg', shape, 'gpu iterations', it, 'numpy', itr, 'relres', rel)
        assert abs(it - itr) <= 1
        assert rel <= tol
        d1, d2 = dp - dp.mean(), xr - xr.mean()
        assert np.abs(d1 - d2).max() <= 1e-8 * np.abs(d2).max()
    solver.close()

"""The MATLAB MEX gateway (matlab/ins2d_mex.c) on the GPU, driven through the
MEX stand-in the way solver_unsteady.m would drive it with
matlab/ins2d_gpu_time_ERK.m: a run through the gateway must give exactly
what the same steps give through the library's C ABI directly (the gateway
only marshals: MATLAB's column-major tableau, cell arrays of samples,
handle table)."""
import numpy as np
import pytest

from ins2d_amd import host
from ins2d_amd.cases import load
from ins2d_amd.rk import get_rk_method
from mex_harness import Cell, Mex, MexError

pytestmark = pytest.mark.gpu

KIND = {'dir': 0, 'per': 1, 'pres': 2, 'sym': 3}


def _say(*a):
    print('[gpu-test]', *a, flush=True)


def _create(mex, case):
    """ins2d_gpu_create.m: the case's grid lines, boundary kinds, Re, tableau
    and solver settings, then the boundary samples of set_bc_vectors.m."""
    solver, opts, V0, p0 = host.setup(case)
    solver.close()
    prm = opts.prm
    bct = case.BCtype()
    bcu = [KIND[bct['u'][s]] for s in ('left', 'right', 'low', 'up')]
    bcv = [KIND[bct['v'][s]] for s in ('left', 'right', 'low', 'up')]
    A, b, c = get_rk_method(prm.get('RK', 'RK44'))
    h = mex('create', opts.grid.x, opts.grid.y, bcu, bcv, float(prm['Re']), int(prm.get('order4', 0)),
            np.asarray(A), b, c, int(prm.get('poisson', 6)), float(prm.get('CG_acc', 1e-8)),
            int(prm.get('CG_maxit', 1000)), int(prm.get('p_add_solve', 1)), int(prm.get('fft_symbol', 0)))
    samples = host.bc_samples(case, opts, opts.t_start)
    mex('set_bc', h, opts.t_start, Cell(samples))
    return h, opts, V0, p0


@pytest.mark.parametrize('name', ['ldc_rk33p2_direct', 'tg_rk44c23_fft'])
def test_gateway_run_equals_direct_run(name):
    _say('matlab gateway run', name)
    mex = Mex()
    if name == 'ldc_rk33p2_direct':
        case = load('LDC_unsteady', RK='RK33P2', poisson=1)
    else:
        # full lower-triangular A: a transposed tableau would change the run
        case = load('taylor_green', Nx=32, Ny=32, Re=300.0, noise=1e-2, RK='RK44C23', dt=0.01, t_end=0.1)
    nsteps = 10
    V_ref, p_ref, rows_ref = host.main(case, nsteps=nsteps)
    h, opts, V0, p0 = _create(mex, case)
    mex('set_state', h, V0, p0)
    mex('input_check', h, opts.t_start, 1)
    t, dt = opts.t_start, opts.dt
    rows = [mex('diagnostics', h, t)]
    for _ in range(nsteps):
        mex('step', h, t, dt, 1)
        t = t + dt
        rows.append(mex('diagnostics', h, t))
    V, p = mex('get_state', h, nout=2)
    nux, nuy, nvx, nvy, npx, npy, xin, yin = mex('grid_info', h, nout=8)
    assert int(nux[0]) * int(nuy[0]) + int(nvx[0]) * int(nvy[0]) == V.size and len(xin) == int(nux[0])
    # the same library calls in the same order: identical results
    _say(name, 'max |V - V_ref|', np.abs(V - V_ref).max(), 'max |p - p_ref|', np.abs(p - p_ref).max())
    assert np.array_equal(V, V_ref) and np.array_equal(p, p_ref)
    assert np.array_equal(np.asarray(rows), np.asarray(rows_ref)[:, 3:])
    # F.m and pressure_poisson.m through the gateway
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(4)
    Vr, pr = rng.standard_normal(V.size), rng.standard_normal(p.size)
    mr, F = mex('rhs', h, Vr, pr, 0.0, nout=2)
    mr_ref, F_ref = solver.rhs(Vr, pr)
    assert np.array_equal(F, F_ref) and mr[0] == mr_ref
    f = rng.standard_normal(p.size)
    f -= f.mean()
    dp = mex('poisson', h, f)
    dp_ref, _, _ = solver.poisson(f)
    assert np.array_equal(dp, dp_ref)
    solver.close()
    mex('destroy', h)
    with pytest.raises(MexError, match='INS2D:handle'):
        mex('step', h, t, dt, 1)


def test_gateway_reports_pcg_nonconvergence():
    """cg.F:262 raises a MATLAB error when PCG does not converge; the
    gateway raises INS2D:noconv with the library's message."""
    mex = Mex()
    case = load('LDC_unsteady', Nx=32, Ny=32, poisson=3, CG_acc=1e-14, CG_maxit=3)
    h, opts, V0, p0 = _create(mex, case)
    f = np.random.default_rng(1).standard_normal(32 * 32)
    f -= f.mean()
    with pytest.raises(MexError, match='INS2D:noconv'):
        mex('poisson', h, f)
    mex('destroy', h)

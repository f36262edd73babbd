"""Pin the CPU oracle by what the mathematics fixes (closed forms, symbols,
invariants and truncation-error rates), for the parts the reference-held
fixtures leave open: the fourth-order operators, the FFT branch, the IC(0)
PCG and the discrete-curl initial field.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed operand or an off-by-one closure) fails it."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from ins2d_amd.cases import load
from oracle import ins2d_oracle as O

WALL = {'BC_unsteady': 0, 'u': dict(left='dir', right='dir', low='dir', up='dir'),
        'v': dict(left='dir', right='dir', low='dir', up='dir')}


def _opts(case='taylor_green', **kw):
    base = dict(Nx=16, Ny=16, noise=0.0, t_end=1.0, dt=0.01)
    base.update(kw)
    return O.build(load(case, **base))


def _periodic_channel(Nx=12, Ny=10, **kw):
    c = load('LDC_unsteady', Nx=Nx, Ny=Ny, **kw)
    c.BCtype = lambda: {'BC_unsteady': 0, 'u': dict(left='per', right='per', low='dir', up='dir'),
                        'v': dict(left='per', right='per', low='dir', up='dir')}
    return c


def _outflow(Nx=12, Ny=8, **kw):
    c = load('LDC_unsteady', Nx=Nx, Ny=Ny, **kw)
    c.BCtype = lambda: {'BC_unsteady': 0, 'u': dict(left='dir', right='pres', low='dir', up='dir'),
                        'v': dict(left='dir', right='sym', low='dir', up='dir')}
    return c


@pytest.mark.parametrize('maker', ['ldc', 'channel', 'periodic', 'stretched'])
def test_pressure_matrix_row_sums_and_gradient(maker):
    if maker == 'ldc':
        opts = O.build(load('LDC_unsteady', Nx=9, Ny=7))
    elif maker == 'channel':
        opts = O.build(_periodic_channel())
    elif maker == 'stretched':
        opts = O.build(load('LDC_unsteady', Nx=10, Ny=8, sx=1.1, sy=1.05))
    else:
        opts = _opts(Nx=10, Ny=6)
    d = opts.discretization
    # operator_divergence.m:295-300: zero row sums without 'pres' boundaries
    assert np.abs(d.A @ np.ones(opts.grid.Np)).max() < 1e-12
    # A symmetric negative semi-definite (A = -M Om^-1 M^T)
    assert abs(d.A - d.A.T).max() < 1e-14
    ev = np.linalg.eigvalsh(d.A.toarray())
    assert ev.max() < 1e-10


def test_fft_symbol_is_spectrum_of_assembled_A():
    opts = _opts(Nx=12, Ny=8, poisson=6)
    g = opts.grid
    A = opts.discretization.A.toarray()
    Ahat = opts.solversettings.Ahat.copy(); Ahat[0, 0] = 0
    for (kx, ky) in [(0, 1), (3, 2), (5, 7), (6, 4)]:
        mode = np.exp(2j * np.pi * (kx * np.arange(g.Npx)[:, None] / g.Npx + ky * np.arange(g.Npy)[None, :] / g.Npy))
        m = mode.ravel(order='F')
        np.testing.assert_allclose(A @ m, -Ahat[kx, ky] * m, atol=1e-12)


def test_fft_solve_satisfies_assembled_system():
    opts = _opts(Nx=16, Ny=12, poisson=6)
    rng = np.random.default_rng(0)
    f = rng.standard_normal(opts.grid.Np); f -= f.mean()
    dp = O.pressure_poisson(f, 0.0, opts)
    r = opts.discretization.A @ dp - f
    assert np.abs(r).max() < 1e-12 * np.abs(f).max()


@pytest.mark.parametrize('maker', ['ldc', 'outflow'])
def test_ic0_pcg_matches_direct_solve(maker):
    case = load('LDC_unsteady', Nx=14, Ny=10, poisson=4) if maker == 'ldc' else _outflow(poisson=4)
    opts = O.build(case)
    A = opts.discretization.A
    rng = np.random.default_rng(1)
    f = rng.standard_normal(opts.grid.Np)
    if maker == 'ldc':
        f -= f.mean()
    x, it, nrm = O.cg_matlab(A, f, 1e-13, np.zeros_like(f), 2000)
    assert nrm <= 1e-13 and it > 3
    xd = O._singular_direct(A)(f)
    if maker == 'ldc':
        x = x - x.mean(); xd = xd - xd.mean()
    np.testing.assert_allclose(x, xd, atol=1e-9 * np.abs(xd).max())
    # IC(0) recurrence D_i = a_ii - a_{i,i-1}^2/D_{i-1} - a_{i,i-Nx}^2/D_{i-Nx}
    D = O.make_cholinc_diag(A)
    Ad = A.toarray(); Nx = opts.grid.Npx
    for i in (0, 1, Nx, Nx + 3, A.shape[0] - 1):
        s = Ad[i, i]
        if i >= 1: s -= Ad[i, i - 1] ** 2 / D[i - 1]
        if i >= Nx: s -= Ad[i, i - Nx] ** 2 / D[i - Nx]
        assert abs(D[i] - s) < 1e-12 * abs(s)


def _conv_matrix(opts, V):
    """check_symmetry.m:32-35 (and its fourth-order alpha*C1 - C3 analogue)."""
    d = opts.discretization; g = opts.grid
    uh, vh = V[:g.Nu], V[g.Nu:]

    def one(s):
        Cux = getattr(d, 'Cux' + s); Cuy = getattr(d, 'Cuy' + s)
        Cvx = getattr(d, 'Cvx' + s); Cvy = getattr(d, 'Cvy' + s)
        Cu = Cux @ sp.diags(getattr(d, 'Iu_ux' + s) @ uh + getattr(d, 'yIu_ux' + s)) @ getattr(d, 'Au_ux' + s) + \
            Cuy @ sp.diags(getattr(d, 'Iv_uy' + s) @ vh + getattr(d, 'yIv_uy' + s)) @ getattr(d, 'Au_uy' + s)
        Cv = Cvx @ sp.diags(getattr(d, 'Iu_vx' + s) @ uh + getattr(d, 'yIu_vx' + s)) @ getattr(d, 'Av_vx' + s) + \
            Cvy @ sp.diags(getattr(d, 'Iv_vy' + s) @ vh + getattr(d, 'yIv_vy' + s)) @ getattr(d, 'Av_vy' + s)
        return Cu, Cv
    Cu, Cv = one('')
    if d.order4:
        Cu3, Cv3 = one('3')
        Cu = d.alfa * Cu - Cu3; Cv = d.alfa * Cv - Cv3
    return Cu, Cv


def test_convection_skew_symmetric_and_energy_neutral():
    """Second order only: with the beta = 9/8 four-point interpolation of the
    convecting velocity (operator_interpolation.m:171-205) the fourth-order
    operator alfa*C1 - C3 is not exactly skew-symmetric (its diagonal is not a
    combination of p-cell divergences), and the reference checks skew-symmetry
    with the second-order operators only (check_symmetry.m:32-35)."""
    order4 = 0
    opts = _opts(Nx=12, Ny=10, order4=order4, noise=0.0)
    V, p, _ = O.input_check(opts)
    rng = np.random.default_rng(3)
    Cu, Cv = _conv_matrix(opts, V)
    assert abs(Cu + Cu.T).max() < 1e-14 * max(1, abs(Cu).max())
    assert abs(Cv + Cv.T).max() < 1e-14 * max(1, abs(Cv).max())
    cu, cv = O.convection(V, V, 0.0, opts)
    np.testing.assert_allclose(Cu @ V[:opts.grid.Nu], cu, atol=1e-13)
    # semi-discrete kinetic energy is conserved by convection
    assert abs(V @ np.r_[cu, cv]) < 1e-12 * np.abs(np.r_[cu, cv]).max()


def test_wall_convection_skew_symmetric():
    opts = O.build(load('LDC_unsteady', Nx=10, Ny=9))
    rng = np.random.default_rng(5)
    # any discretely divergence-free field with homogeneous walls is a
    # projection of a random field
    V = rng.standard_normal(opts.grid.NV)
    d = opts.discretization
    A = d.A.toarray()
    f = d.M @ V
    dp = np.linalg.lstsq(A, f, rcond=None)[0]
    V = V - opts.grid.Om_inv * (d.G @ dp)
    assert np.abs(d.M @ V).max() < 1e-10
    Cu, Cv = _conv_matrix(opts, V)
    assert abs(Cu + Cu.T).max() < 1e-13


@pytest.mark.parametrize('order4,rate', [(0, 1.9), (1, 3.8)])
def test_taylor_green_truncation_rate(order4, rate):
    """Point-value truncation error of F on the exact Taylor-Green field:
    -(u.grad)u + nu lap u = -sin(2x)/2 - 2 nu sin x cos y at the u points."""
    Re = 3.0
    errs = []
    for N in (16, 32, 64):
        opts = _opts(Nx=N, Ny=N, order4=order4, Re=Re)
        g = opts.grid
        V = opts.V_start
        _, Fr = O.F(V, V, np.zeros(g.Np), 0.0, opts, nopressure=True)
        Fu = Fr[:g.Nu] / g.Omu
        ex = (-0.5 * np.sin(2 * g.xu) - 2 / Re * np.sin(g.xu) * np.cos(g.yu)).ravel(order='F')
        Fv = Fr[g.Nu:] / g.Omv
        exv = (-0.5 * np.sin(2 * g.yv) + 2 / Re * np.cos(g.xv) * np.sin(g.yv)).ravel(order='F')
        errs.append(max(np.abs(Fu - ex).max(), np.abs(Fv - exv).max()))
    r = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(r) > rate, (errs, r)


def test_streamfunction_field_is_divergence_free():
    c = load('decaying_turbulence', Nx=32, Ny=32, order4=0)
    opts = O.build(c)
    d = opts.discretization
    V = opts.V_start
    assert np.abs(d.M @ V + d.yM).max() < 1e-12 * np.abs(V).max() / opts.grid.hx[0]
    g = opts.grid
    e = 0.5 * (np.mean(V[:g.Nu] ** 2) + np.mean(V[g.Nu:] ** 2))
    assert abs(e - 1.0) < 1e-12


def test_inviscid_periodic_momentum_conserved_and_energy_drift_small():
    c = load('shear_layer', Nx=24, Ny=24, RK='RK44', dt=0.01, t_end=0.2)
    _, _, rows, _ = O.run(c)
    rows = np.asarray(rows)
    assert np.abs(rows[:, 5] - rows[0, 5]).max() < 1e-12
    assert np.ptp(rows[:, 7]) / rows[0, 7] < 1e-9


def test_fourth_order_projection_gives_divergence_free_field():
    opts = _opts(Nx=12, Ny=12, order4=1, noise=1e-2, poisson=1)
    V, p, row = O.input_check(opts)
    d = opts.discretization
    assert np.abs(d.M @ V + d.yM).max() < 1e-12


@pytest.mark.parametrize('maker', ['ldc', 'outflow', 'channel'])
def test_ilu0_factor_and_pcg(maker):
    """poisson = 2 (pressure_poisson.m:52-64 with operator_divergence.m:243-247):
    the oracle's ILU(0) reproduces A on its pattern (the defining property
    of a no-fill incomplete LU), is the transpose-symmetric factor of a
    symmetric matrix, and its pivots are make_cholinc.m's IC(0) diagonal (two
    restatements: row-by-row elimination against the diagonal-storage
    recurrence); its PCG agrees with SciPy's CG given the same
    preconditioner M = L L^T."""
    import scipy.sparse.linalg as spla
    if maker == 'ldc':
        opts = O.build(load('LDC_unsteady', Nx=14, Ny=10, poisson=2, CG_acc=1e-10, CG_maxit=2000))
    elif maker == 'outflow':
        opts = O.build(_outflow(poisson=2, CG_acc=1e-10, CG_maxit=2000))
    else:
        opts = O.build(_periodic_channel(poisson=2, CG_acc=1e-10, CG_maxit=2000))
    A = opts.discretization.A.tocsr()
    L, U = opts.solversettings.L, opts.solversettings.U
    LU = (L @ U).tocsr()
    coo = A.tocoo()
    assert np.abs(np.asarray(LU[coo.row, coo.col]).ravel() - coo.data).max() < 1e-12 * np.abs(coo.data).max()
    assert abs(U - sp.diags(U.diagonal()) @ L.T).max() < 1e-12 * np.abs(coo.data).max()
    np.testing.assert_allclose(U.diagonal(), O.make_cholinc_diag(A), rtol=1e-13)
    rng = np.random.default_rng(2)
    f = rng.standard_normal(opts.grid.Np)
    if maker != 'outflow':
        f -= f.mean()
    x, it = O.pcg_ilu(A, f, 1e-10, 2000, L)
    K = (-A).tocsr()
    Lt = L.T.tocsr()
    Minv = spla.LinearOperator(K.shape, matvec=lambda r: spla.spsolve_triangular(
        Lt, spla.spsolve_triangular(L, r, lower=True), lower=False))
    count = []
    xs, info = spla.cg(K, -f, rtol=1e-10, atol=0.0, maxiter=2000, M=Minv, callback=lambda xk: count.append(1))
    assert info == 0 and abs(len(count) - it) <= 1
    np.testing.assert_allclose(x, xs, atol=1e-7 * np.abs(xs).max())
    assert np.linalg.norm(A @ x - f) <= 1.01e-10 * np.linalg.norm(f)

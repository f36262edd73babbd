"""Pin the oracle's 'pres' outflow closures and its stretched-grid weights by
what the mathematics fixes, independently of the reference-held fixtures
(which are all uniform, closed or periodic):

* a parallel flow u(y), v = 0 through a channel with a 'pres' outlet: every
  column of Om^-1 F must be the same, the outlet column included, and the
  v-equation must vanish -- the outlet closures of the averaging, the
  diffusion and the half-cell interpolation (spatial/operators/
  operator_interpolation.m:82-90,122-130) all enter the outlet column;
* a smooth manufactured field with v != 0 and zero normal gradient at the
  outlet: the point error of Om^-1 F against the continuous
  -div(u u) + lap(u)/Re converges under refinement at every u and v point,
  the outlet column included, on uniform, geometric (libs/nonuniform_grid.m)
  and tanh-stretched grids -- the mat_hx2 / mat_hy2 flux weights
  (operator_interpolation.m:49-58) and the stretched diffusion weights
  (operator_convection_diffusion.m:133-161) enter every point;
* skew-symmetry of the convection operator on stretched wall-bounded grids
  for a discretely divergence-free field (spatial/check_symmetry.m:42-53).

Each check fails for a plausible restatement mistake: the half-cell factor
at the outlet dropped (the outlet column's error stays O(1)), a stretched
weight taken from the wrong neighbour (first-order error turns O(1) or the
skew-symmetry breaks), a wrong sign in a closure.  The mistakes are planted
in copies of the oracle's operators below, on the CPU, to show that the
checks see them."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from ins2d_amd.cases import load
from ins2d_amd.cases._common import nonuniform_grid, tanh_mesh_1d
from oracle import ins2d_oracle as O

L = 2.0


def _grid_lines(kind, a, b, n, s):
    if kind == 'uniform':
        return np.linspace(a, b, n + 1)
    if kind == 'geometric':
        # libs/nonuniform_grid.m with a stretch factor per cell chosen so that
        # the overall ratio of last to first cell stays fixed under refinement
        sz = s ** (16.0 / n)
        # the routine counts the cells of first spacing d0 that fit and then
        # refits the factor to end at b: shrink d0 until n cells fit
        d0 = (b - a) * (sz - 1.0) / (sz ** n - 1.0)
        for _ in range(200):
            z = nonuniform_grid(d0, a, b, sz)
            if len(z) >= n + 1:
                break
            d0 *= 0.995
        assert len(z) == n + 1 and abs(z[-1] - b) < 1e-12
        return z
    return tanh_mesh_1d(a, b, n, s)


def _fields(outlet):
    """Smooth u, v with zero normal gradient at the outlet (x = L for the
    right outlet, y = 1/2 for the top outlet), and the continuous
    momentum right-hand sides -div(u u) + lap(u)/Re for each component."""
    import sympy as S
    x, y, Re = S.symbols('x y Re')
    if outlet == 'right':
        # d/dx of u and v vanish at x = L
        u = 1 + S.Rational(3, 10) * S.cos(S.pi * x / L) * S.sin(S.pi * y + S.Rational(3, 10))
        v = S.Rational(1, 5) * S.cos(S.pi * x / L) * S.cos(2 * S.pi * y + S.Rational(1, 10))
    else:
        # d/dy of u and v vanish at y = 1/2 (domain y in [-1/2, 1/2])
        u = S.Rational(1, 5) * S.cos(S.pi * (y - S.Rational(1, 2))) * S.cos(2 * S.pi * x / L + S.Rational(1, 10))
        v = 1 + S.Rational(3, 10) * S.cos(S.pi * (y - S.Rational(1, 2))) * S.sin(S.pi * x / L + S.Rational(3, 10))
    fu = -(S.diff(u * u, x) + S.diff(u * v, y)) + (S.diff(u, x, 2) + S.diff(u, y, 2)) / Re
    fv = -(S.diff(u * v, x) + S.diff(v * v, y)) + (S.diff(v, x, 2) + S.diff(v, y, 2)) / Re
    lam = lambda e: S.lambdify((x, y, Re), e, 'numpy')
    return lam(u), lam(v), lam(fu), lam(fv)


def _case(outlet, xkind, ykind, Nx, Ny, Re=50.0, fields=None, parallel=False):
    """A channel case on [0, L] x [-1/2, 1/2] with one 'pres' outlet and
    Dirichlet values of the exact field on the other sides."""
    c = load('LDC_unsteady', Nx=Nx, Ny=Ny, Re=Re, x1=0.0, x2=L, y1=-0.5, y2=0.5)
    if outlet == 'right':
        bct = {'BC_unsteady': 0, 'u': dict(left='dir', right='pres', low='dir', up='dir'),
               'v': dict(left='dir', right='sym', low='dir', up='dir')}
    else:
        bct = {'BC_unsteady': 0, 'u': dict(left='dir', right='dir', low='dir', up='sym'),
               'v': dict(left='dir', right='dir', low='dir', up='pres')}
    c.BCtype = lambda: bct
    c.mesh = lambda opts: (_grid_lines(xkind, 0.0, L, Nx, 1.15), _grid_lines(ykind, -0.5, 0.5, Ny, 1.2))
    if parallel:
        # u(y) = 1 - 4 y^2 + y / 3, v = 0
        uf = lambda X, Y: (1.0 - 4.0 * Y ** 2 + Y / 3.0) + 0.0 * X
        vf = lambda X, Y: 0.0 * X * Y
    else:
        uf = lambda X, Y: fields[0](X, Y, Re) + 0.0 * X
        vf = lambda X, Y: fields[1](X, Y, Re) + 0.0 * X

    def sampler(fun, normal_zero):
        def bc(xs, ys, t, options):
            xs = np.atleast_1d(np.asarray(xs, float)); ys = np.atleast_1d(np.asarray(ys, float))
            # Neumann sides ('pres' / 'sym'): the sample is the normal
            # derivative (BC_general.m:49-52), zero for these fields
            if outlet == 'right' and len(xs) == 1 and abs(xs[0] - L) < 1e-12 and normal_zero:
                return np.zeros(len(ys))
            if outlet == 'top' and len(ys) == 1 and abs(ys[0] - 0.5) < 1e-12 and normal_zero:
                return np.zeros(len(xs))
            X, Y = np.meshgrid(xs, ys, indexing='ij')
            return np.asarray(fun(X, Y), float).ravel(order='F')
        return bc

    c.uBC = sampler(uf, True)
    c.vBC = sampler(vf, True)

    def IC(t, options):
        g = options.grid
        return uf(g.xu, g.yu), vf(g.xv, g.yv), np.zeros((g.Npx, g.Npy))
    c.IC = IC
    return c


def _rhs(opts):
    g = opts.grid
    _, Fr = O.F(opts.V_start, opts.V_start, np.zeros(g.Np), 0.0, opts, nopressure=True)
    return Fr[:g.Nu] / g.Omu, Fr[g.Nu:] / g.Omv


@pytest.mark.parametrize('xkind,ykind', [('uniform', 'uniform'), ('geometric', 'tanh'), ('tanh', 'geometric')])
def test_parallel_channel_flow_is_column_invariant(xkind, ykind):
    opts = O.build(_case('right', xkind, ykind, 12, 10, parallel=True))
    g = opts.grid
    Fu, Fv = _rhs(opts)
    Fu = Fu.reshape(g.Nux_in, g.Nuy_in, order='F')
    # the outlet keeps its face: Nux_in = Nx (operator_mesh.m:54-98)
    assert g.Nux_in == g.Nx
    scale = np.abs(Fu).max()
    assert scale > 1e-3
    # Om^-1 F identical in every column, the outlet column included
    assert np.abs(Fu - Fu[:1, :]).max() < 1e-12 * scale
    # and equal to the y-diffusion alone: with u = u(y), v = 0 there is no
    # convection and no x-diffusion
    assert np.abs(Fv).max() < 1e-12 * scale


def _errors(outlet, xkind, ykind, N):
    fields = _fields(outlet)
    Re = 50.0
    Nx, Ny = (2 * N, N) if outlet == 'right' else (N, 2 * N)
    opts = O.build(_case(outlet, xkind, ykind, Nx, Ny, Re=Re, fields=fields))
    g = opts.grid
    Fu, Fv = _rhs(opts)
    eu = np.abs(Fu - fields[2](g.xu, g.yu, Re).ravel(order='F')).reshape(g.Nux_in, g.Nuy_in, order='F')
    ev = np.abs(Fv - fields[3](g.xv, g.yv, Re).ravel(order='F')).reshape(g.Nvx_in, g.Nvy_in, order='F')
    if outlet == 'right':
        out = max(eu[-1, :].max(), ev[-1, :].max())
    else:
        out = max(eu[:, -1].max(), ev[:, -1].max())
    return max(eu.max(), ev.max()), out


@pytest.mark.parametrize('outlet', ['right', 'top'])
@pytest.mark.parametrize('xkind,ykind', [('uniform', 'uniform'), ('geometric', 'tanh'), ('tanh', 'geometric')])
def test_manufactured_field_truncation_converges_at_outlet(outlet, xkind, ykind):
    """First order at least everywhere (the staggered control volumes of a
    stretched or half cell are not centred on their velocity point), the
    outlet column included."""
    res = [_errors(outlet, xkind, ykind, N) for N in (8, 16, 32, 64)]
    all_e = [r[0] for r in res]
    out_e = [r[1] for r in res]
    r_all = [math.log2(all_e[i] / all_e[i + 1]) for i in range(3)]
    r_out = [math.log2(out_e[i] / out_e[i + 1]) for i in range(3)]
    assert min(r_all[1:]) > 0.85, (all_e, r_all)
    assert min(r_out[1:]) > 0.85, (out_e, r_out)
    assert all_e[-1] < 0.2 * all_e[0]


def test_outlet_half_cell_factor_is_seen():
    """The outlet column's error of the manufactured field stays O(1) when
    the 2022 half-cell factor of Iv_uy is undone (a planted mistake in a copy
    of the oracle's operator), so the convergence test above pins it."""
    fields = _fields('right')
    errs = []
    for N in (16, 32):
        opts = O.build(_case('right', 'uniform', 'uniform', 2 * N, N, fields=fields))
        g = opts.grid
        d = opts.discretization
        # Iv_uy's outlet row block is the last of every Nux_in rows: undo the
        # halving by doubling those rows (operator and boundary vector)
        rows = np.arange(g.Nux_in - 1, d.Iv_uy.shape[0], g.Nux_in)
        s = np.ones(d.Iv_uy.shape[0]); s[rows] = 2.0
        d.Iv_uy = sp.diags(s) @ d.Iv_uy
        d.yIv_uy = s * d.yIv_uy
        Fu, _ = _rhs(opts)
        eu = np.abs(Fu - fields[2](g.xu, g.yu, 50.0).ravel(order='F')).reshape(g.Nux_in, g.Nuy_in, order='F')
        errs.append(eu[-1, :].max())
    assert errs[1] > 0.5 * errs[0] and errs[1] > 0.05, errs


# ---------------------------------------------------------------------------
# exactness on geometric grids of fixed cell ratio (nonuniform_grid.m with a
# fixed stretch factor): the cell widths jump by the factor from cell to cell
# at every resolution, so a weight taken from the wrong cell is an O(1) error,
# while fields of the kinds below are differentiated exactly by the correct
# operators (linear profiles across the stretched direction are averaged to
# the faces exactly because every face lies midway between its neighbours in
# the other direction, and every flux weight is the face length)


def _exact_case(outlet, xs, ys, uf, vf, dn_u=None, dn_v=None, Re=7.0, x2=L):
    """A channel with Dirichlet values of the field on every side but the
    outlet ('right': u 'pres', v 'sym'; 'top': v 'pres', u 'sym'; None: all
    'dir'), where the sample is the field's normal derivative dn_u / dn_v
    (the Neumann value of BC_general.m:49-52)."""
    Nx, Ny = len(xs) - 1, len(ys) - 1
    c = load('LDC_unsteady', Nx=Nx, Ny=Ny, Re=Re, x1=xs[0], x2=xs[-1], y1=ys[0], y2=ys[-1])
    u_bc = dict(left='dir', right='dir', low='dir', up='dir')
    v_bc = dict(left='dir', right='dir', low='dir', up='dir')
    if outlet == 'right':
        u_bc['right'] = 'pres'; v_bc['right'] = 'sym'
    elif outlet == 'top':
        u_bc['up'] = 'sym'; v_bc['up'] = 'pres'
    c.BCtype = lambda: {'BC_unsteady': 0, 'u': u_bc, 'v': v_bc}
    c.mesh = lambda opts: (np.asarray(xs), np.asarray(ys))

    def sampler(fun, dn):
        def bc(x, y, t, options):
            x = np.atleast_1d(np.asarray(x, float)); y = np.atleast_1d(np.asarray(y, float))
            X, Y = np.meshgrid(x, y, indexing='ij')
            if outlet == 'right' and len(x) == 1 and abs(x[0] - xs[-1]) < 1e-12:
                return np.asarray(dn(X, Y) + 0.0 * X, float).ravel(order='F')
            if outlet == 'top' and len(y) == 1 and abs(y[0] - ys[-1]) < 1e-12:
                return np.asarray(dn(X, Y) + 0.0 * X, float).ravel(order='F')
            return np.asarray(fun(X, Y) + 0.0 * X, float).ravel(order='F')
        return bc

    zero = lambda X, Y: 0.0 * X
    c.uBC = sampler(uf, dn_u or zero)
    c.vBC = sampler(vf, dn_v or zero)
    c.IC = lambda t, o: (uf(o.grid.xu, o.grid.yu) + 0.0 * o.grid.xu, vf(o.grid.xv, o.grid.yv) + 0.0 * o.grid.xv,
                         np.zeros((o.grid.Npx, o.grid.Npy)))
    return c


def _geo(a, b, sz, d0):
    z = nonuniform_grid(d0, a, b, sz)
    h = np.diff(z)
    assert h.max() / h.min() > 2.0
    return z


XS = _geo(0.0, L, 1.25, 0.06)           # the cell width grows by 1.25 per cell
YS = _geo(-0.5, 0.5, 1.3, 0.03)
XU = np.linspace(0.0, L, 13)
YU = np.linspace(-0.5, 0.5, 9)


def _exact_checks():
    """(case, expected Om^-1 F_u, expected Om^-1 F_v) with the continuous
    values -div(u u) + lap(u)/Re of the fields."""
    a, b, V0, U0 = 0.7, 0.9, 0.4, -0.3
    return {
        # x stretched: u = a + b y, v = V0 -> F_u = -V0 b, F_v = 0; the outlet
        # keeps its half cell (Iv_uy's mat_hx2 weights and the half-cell
        # factor of operator_interpolation.m:82-90)
        'x_stretched_uy_right_outlet': (
            _exact_case('right', XS, YU, lambda X, Y: a + b * Y, lambda X, Y: V0 + 0.0 * X),
            lambda X, Y: -V0 * b + 0.0 * X, lambda X, Y: 0.0 * X),
        # y stretched: v = a + b x, u = U0 -> F_v = -U0 b, F_u = 0 (Iu_vx's
        # mat_hy2 weights and operator_interpolation.m:122-130)
        'y_stretched_vx_top_outlet': (
            _exact_case('top', XU, YS, lambda X, Y: U0 + 0.0 * X, lambda X, Y: a + b * X),
            lambda X, Y: 0.0 * X, lambda X, Y: -U0 * b + 0.0 * X),
        # y stretched: u = a + b x, v = 0 -> F_u = -2 b u (Iu_ux's mat_hy)
        'y_stretched_ux_closed': (
            _exact_case(None, XU, YS, lambda X, Y: a + b * X, lambda X, Y: 0.0 * X),
            lambda X, Y: -2 * b * (a + b * X), lambda X, Y: 0.0 * X),
        # x stretched: v = a + b y, u = 0 -> F_v = -2 b v (Iv_vy's mat_hx)
        'x_stretched_vy_closed': (
            _exact_case(None, XS, YU, lambda X, Y: 0.0 * X, lambda X, Y: a + b * Y),
            lambda X, Y: 0.0 * X, lambda X, Y: -2 * b * (a + b * Y)),
    }


@pytest.mark.parametrize('name', ['x_stretched_uy_right_outlet', 'y_stretched_vx_top_outlet',
                                  'y_stretched_ux_closed', 'x_stretched_vy_closed'])
def test_exact_on_fixed_ratio_stretched_grids(name):
    case, eu, ev = _exact_checks()[name]
    opts = O.build(case)
    g = opts.grid
    Fu, Fv = _rhs(opts)
    ex_u = (eu(g.xu, g.yu) + 0.0 * g.xu).ravel(order='F')
    ex_v = (ev(g.xv, g.yv) + 0.0 * g.xv).ravel(order='F')
    sc = max(np.abs(ex_u).max(), np.abs(ex_v).max(), 1.0)
    assert np.abs(Fu - ex_u).max() < 1e-12 * sc, np.abs(Fu - ex_u).max()
    assert np.abs(Fv - ex_v).max() < 1e-12 * sc, np.abs(Fv - ex_v).max()


def test_diffusion_exact_for_linear_fields_on_stretched_grids():
    """lap of a linear field is zero: the stretched diffusion weights
    (operator_convection_diffusion.m:133-161) and the Dirichlet, staggered
    mirror and 'pres'/'sym' closures give zero at every point.  At an outlet
    the field has zero normal derivative, the value every reference case
    samples there: with a nonzero sample the 'pres' closure puts the ghost
    at u_N + 2 h du (BC_general.m:49-52,66-69), twice the one-sided
    difference the sample would suggest -- a quirk no case exercises."""
    checks = [
        (None, lambda X, Y: 0.3 + 0.8 * X - 0.6 * Y, lambda X, Y: -0.2 + 0.5 * X + 0.9 * Y),
        ('right', lambda X, Y: 0.3 - 0.6 * Y + 0.0 * X, lambda X, Y: -0.2 + 0.9 * Y + 0.0 * X),
        ('top', lambda X, Y: 0.3 + 0.8 * X + 0.0 * Y, lambda X, Y: -0.2 + 0.5 * X + 0.0 * Y),
    ]
    for outlet, uf, vf in checks:
        opts = O.build(_exact_case(outlet, XS, YS, uf, vf))
        d2u, d2v = O.diffusion(opts.V_start, 0.0, opts)
        g = opts.grid
        assert np.abs(d2u / g.Omu).max() < 1e-11, (outlet, np.abs(d2u / g.Omu).max())
        assert np.abs(d2v / g.Omv).max() < 1e-11, (outlet, np.abs(d2v / g.Omv).max())
    # a stretched diffusion weight from the neighbouring cell is seen
    opts = O.build(_exact_case(None, XS, YS, checks[0][1], checks[0][2]))
    d = opts.discretization
    g = opts.grid
    hx = np.asarray(g.hx)
    rows = np.arange(d.Su_ux.shape[0]) % (g.Nux_in + 1)
    s = np.r_[hx[1:] / hx[:-1], 1.0][np.minimum(rows, len(hx) - 1)]
    d.Diffu = d.Dux @ sp.diags(s / opts.fluid.Re) @ d.Su_ux + d.Duy @ ((1 / opts.fluid.Re) * d.Su_uy)
    d2u, _ = O.diffusion(opts.V_start, 0.0, opts)
    assert np.abs(d2u / g.Omu).max() > 1e-2


def test_planted_weight_mistakes_are_seen():
    """Mistakes planted in copies of the oracle's operators (on the CPU) fail
    the exactness checks: the half-cell factor at the outlet undone, and the
    flux weights of Iv_uy taken from the neighbouring cell."""
    case, eu, ev = _exact_checks()['x_stretched_uy_right_outlet']
    for mistake in ('half_cell', 'neighbour'):
        opts = O.build(case)
        g = opts.grid
        d = opts.discretization
        rows = np.arange(d.Iv_uy.shape[0]) % g.Nux_in
        if mistake == 'half_cell':
            s = np.where(rows == g.Nux_in - 1, 2.0, 1.0)
        else:
            hx = np.asarray(g.hx)
            s = np.r_[hx[1:] / hx[:-1], 1.0][rows]
        d.Iv_uy = sp.diags(s) @ d.Iv_uy
        d.yIv_uy = s * d.yIv_uy
        Fu, _ = _rhs(opts)
        ex_u = (eu(g.xu, g.yu) + 0.0 * g.xu).ravel(order='F')
        assert np.abs(Fu - ex_u).max() > 1e-2, mistake


def _streamfunction_field(opts, seed):
    """u = (psi_{i,j+1} - psi_{i,j}) / hy_j, v = -(psi_{i+1,j} - psi_{i,j}) /
    hx_i from a random corner streamfunction that vanishes on the walls:
    M V = 0 exactly (the face fluxes telescope) and no flow through walls."""
    g = opts.grid
    rng = np.random.default_rng(seed)
    psi = np.zeros((g.Nx + 1, g.Ny + 1))
    psi[1:-1, 1:-1] = rng.standard_normal((g.Nx - 1, g.Ny - 1))
    hx = np.asarray(g.hx); hy = np.asarray(g.hy)
    u = (psi[1:-1, 1:] - psi[1:-1, :-1]) / hy[None, :]
    v = -(psi[1:, 1:-1] - psi[:-1, 1:-1]) / hx[:, None]
    return np.r_[u.ravel(order='F'), v.ravel(order='F')]


@pytest.mark.parametrize('xkind,ykind', [('geometric', 'tanh'), ('tanh', 'geometric'), ('tanh', 'tanh')])
def test_stretched_wall_convection_skew_symmetric(xkind, ykind):
    c = load('LDC_unsteady', Nx=11, Ny=9)
    c.mesh = lambda opts: (_grid_lines(xkind, 0.0, 1.0, 11, 1.3), _grid_lines(ykind, 0.0, 1.0, 9, 1.4))
    c.uBC = lambda x, y, t, o: np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))
    opts = O.build(c)
    g = opts.grid
    V = _streamfunction_field(opts, 7)
    d = opts.discretization
    assert np.abs(d.M @ V + d.yM).max() < 1e-12 * np.abs(V).max()
    from tests.test_oracle_invariants import _conv_matrix
    Cu, Cv = _conv_matrix(opts, V)
    sc = max(abs(Cu).max(), abs(Cv).max())
    assert abs(Cu + Cu.T).max() < 1e-14 * sc
    assert abs(Cv + Cv.T).max() < 1e-14 * sc
    # a weight from the wrong neighbour breaks it
    hx = np.asarray(g.hx)
    s = np.r_[hx[1:] / hx[:-1], 1.0][np.arange(d.Iv_uy.shape[0]) % g.Nux_in]
    d.Iv_uy = sp.diags(s) @ d.Iv_uy
    Cu2, _ = _conv_matrix(opts, V)
    assert abs(Cu2 + Cu2.T).max() > 1e-6 * sc

"""Pins of the oracle's fourth-order operators on grids with walls
(operator_*.m with order4 = 1 and Dirichlet or periodic ends): properties
the reference's discretization has by construction, each checked so that a
plausible mistake in a closure, a boundary vector or a weight fails one of
them.

  * constant fields: zero convection and divergence (the averaging and
    interpolation closures and their boundary vectors reproduce constants);
  * linear fields: zero diffusion; a linear divergence-free field has zero
    discrete divergence (the 3h divergence closure reproduces linears);
  * energy: V . conv(V) = 0 for a discretely divergence-free V with zero wall
    velocities (the skew-symmetric convection of the energy-conserving
    scheme, at walls too);
  * structure: G = -M^T, A = M Om^-1 G symmetric with the constant null
    vector and no positive eigenvalue, diffusion with real non-positive
    spectrum;
  * interior accuracy: fourth order away from the walls (16 -> 32 -> 64
    cells), and a planted mistake (beta = 1, the second-order
    interpolation weights) drops the rate to second order.

Near the walls the reference's closures are first order or inconsistent for
general fields (O(1) truncation at the first points); that is the
reference's design and is not tested here beyond the exactness above."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ins2d_oracle as O
from w4_cases import wall_case


def _case(ufun, vfun, pery=False, N=(14, 12), Re=200.0, uniform=False):
    c = wall_case(pery, N=N, Re=Re)
    grid = lambda a: np.meshgrid(np.atleast_1d(a[0]), np.atleast_1d(a[1]), indexing='ij')
    c.uBC = lambda x, y, t, o: ufun(*grid((x, y))).ravel(order='F')
    c.vBC = lambda x, y, t, o: vfun(*grid((x, y))).ravel(order='F')
    if uniform:
        c.mesh = lambda opts: (np.linspace(0, 1, opts.grid.Nx + 1), np.linspace(0, 1, opts.grid.Ny + 1))
    return c


def _field(opts, ufun, vfun):
    g = opts.grid
    return np.r_[ufun(g.xu, g.yu).ravel(order='F'), vfun(g.xv, g.yv).ravel(order='F')]


@pytest.mark.parametrize('pery', [False, True])
def test_constant_field_has_no_convection_or_divergence(pery):
    uf = lambda X, Y: 0.7 + 0 * X
    vf = lambda X, Y: -0.4 + 0 * X
    o = O.build(_case(uf, vf, pery))
    V = _field(o, uf, vf)
    cu, cv = O.convection(V, V, 0.0, o)
    assert np.abs(np.r_[cu, cv]).max() < 1e-13
    d = o.discretization
    assert np.abs(d.M @ V + d.yM).max() < 1e-13


@pytest.mark.parametrize('pery', [False, True])
def test_linear_fields(pery):
    if pery:   # linear in x only (periodic in y)
        uf = lambda X, Y: 0.3 + 0 * X
        vf = lambda X, Y: -0.2 + 0.3 * X
    else:
        uf = lambda X, Y: 0.3 + 0.5 * X + 0.8 * Y
        vf = lambda X, Y: -0.2 - 0.5 * Y + 0.3 * X
    o = O.build(_case(uf, vf, pery, Re=1.0))
    V = _field(o, uf, vf)
    du, dv = O.diffusion(V, 0.0, o)
    d = o.discretization
    # relative to the size of the terms that cancel (row sums of |Diff| |V|)
    scale = max(abs(d.Diffu).sum(axis=1).max(), abs(d.Diffv).sum(axis=1).max()) * np.abs(V).max()
    assert np.abs(np.r_[du, dv]).max() < 1e-14 * scale
    assert np.abs(d.M @ V + d.yM).max() < 1e-13   # divergence-free linear field


@pytest.mark.parametrize('pery', [False, True])
def test_convection_conserves_energy(pery):
    zf = lambda X, Y: 0 * X
    o = O.build(_case(zf, zf, pery, Re=1e100))
    d, g = o.discretization, o.grid
    rng = np.random.default_rng(1)
    V = rng.standard_normal(g.NV)
    V = V - g.Om_inv * (d.G @ o.solversettings.decomp(d.M @ V))
    assert np.abs(d.M @ V).max() < 1e-12
    cu, cv = O.convection(V, V, 0.0, o)
    c = np.r_[cu, cv]
    assert abs(V @ c) < 1e-14 * np.abs(V).max() * np.abs(c).max() * len(V)
    if not pery:
        # a planted mistake: the 3h y-difference of u without its one-sided
        # rows at the walls (operator_convection_diffusion.m:300-315) breaks
        # the energy balance
        D13 = O.mspdiags([-np.ones(g.Nuy_t + 5), np.ones(g.Nuy_t + 5)], [0, 3], g.Nuy_t - 2, g.Nuy_t + 1)
        d.Cuy3 = O.kron(D13, O.speye(g.Nux_in))
        bad = np.r_[O.convection(V, V, 0.0, o)]
        assert abs(V @ bad.ravel()) > 1e-6 * np.abs(V).max() * np.abs(c).max()


@pytest.mark.parametrize('pery', [False, True])
def test_operator_structure(pery):
    zf = lambda X, Y: 0 * X
    o = O.build(_case(zf, zf, pery, Re=1.0))
    d = o.discretization
    assert abs(d.G + d.M.T).max() == 0.0
    A = d.A.toarray()
    scale = np.abs(A).max()
    assert np.abs(A - A.T).max() < 1e-14 * scale
    assert np.abs(A.sum(axis=1)).max() < 1e-13 * scale
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 1e-12 * scale
    D = sp.block_diag([d.Diffu, d.Diffv]).toarray()
    ev = np.linalg.eigvals(D)
    assert np.abs(ev.imag).max() < 1e-10 * np.abs(ev).max()
    assert ev.real.max() < 1e-12 * np.abs(ev).max()


def _interior_errors(N, beta=None):
    a, b, c, dd, e, f, gg, k = 1.3, 0.4, 0.9, 0.2, 0.7, 0.1, 1.1, 0.3
    uf = lambda X, Y: np.sin(a * X + b) * np.cos(c * Y + dd)
    vf = lambda X, Y: np.cos(e * X + f) * np.sin(gg * Y + k)

    def conv_u(X, Y):   # d(uu)/dx + d(vu)/dy
        U, V = uf(X, Y), vf(X, Y)
        Ux, Uy = a * np.cos(a * X + b) * np.cos(c * Y + dd), -c * np.sin(a * X + b) * np.sin(c * Y + dd)
        Vy = gg * np.cos(e * X + f) * np.cos(gg * Y + k)
        return 2 * U * Ux + Vy * U + V * Uy

    def conv_v(X, Y):   # d(uv)/dx + d(vv)/dy
        U, V = uf(X, Y), vf(X, Y)
        Ux = a * np.cos(a * X + b) * np.cos(c * Y + dd)
        Vx, Vy = -e * np.sin(e * X + f) * np.sin(gg * Y + k), gg * np.cos(e * X + f) * np.cos(gg * Y + k)
        return Ux * V + U * Vx + 2 * V * Vy

    lap_u = lambda X, Y: -(a * a + c * c) * uf(X, Y)
    overrides = {'beta': beta} if beta is not None else None
    o = O.build(_case(uf, vf, False, N=(N, N), Re=1.0, uniform=True), overrides)
    g = o.grid
    V = _field(o, uf, vf)
    Cu, Cv = O.convection(V, V, 0.0, o)
    Du, _ = O.diffusion(V, 0.0, o)
    h = 1.0 / N
    inner = lambda X, Y: ((X > 4 * h - 1e-12) & (X < 1 - 4 * h + 1e-12) & (Y > 4 * h - 1e-12) &
                          (Y < 1 - 4 * h + 1e-12)).ravel(order='F')
    mu, mv = inner(g.xu, g.yu), inner(g.xv, g.yv)
    ecu = np.abs(Cu / g.Omu - conv_u(g.xu, g.yu).ravel(order='F'))[mu].max()
    ecv = np.abs(Cv / g.Omv - conv_v(g.xv, g.yv).ravel(order='F'))[mv].max()
    edu = np.abs(Du / g.Omu - lap_u(g.xu, g.yu).ravel(order='F'))[mu].max()
    return np.array([ecu, ecv, edu])


def test_interior_fourth_order():
    e16, e32, e64 = (_interior_errors(N) for N in (16, 32, 64))
    r1, r2 = np.log2(e16 / e32), np.log2(e32 / e64)
    assert np.all(r1 > 3.5) and np.all(r2 > 3.7), (r1, r2)
    # planted mistake: second-order interpolation weights (beta = 1)
    b32, b64 = _interior_errors(32, beta=1.0), _interior_errors(64, beta=1.0)
    rb = np.log2(b32 / b64)
    assert rb[0] < 2.5 and rb[1] < 2.5, rb


def test_convection_conserves_energy_fully_periodic():
    """The same energy balance on the periodic fourth-order operators."""
    from ins2d_amd.cases import load
    o = O.build(load('taylor_green', Nx=16, Ny=16, order4=1, poisson=1, Re=1e100))
    d, g = o.discretization, o.grid
    rng = np.random.default_rng(1)
    V = rng.standard_normal(g.NV)
    V = V - g.Om_inv * (d.G @ o.solversettings.decomp(d.M @ V))
    assert np.abs(d.M @ V).max() < 1e-12
    c = np.r_[O.convection(V, V, 0.0, o)].ravel()
    assert abs(V @ c) < 1e-14 * np.abs(V).max() * np.abs(c).max() * len(V)

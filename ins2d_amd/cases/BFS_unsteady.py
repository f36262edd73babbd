"""Backward-facing step channel (case_files/BFS_unsteady/BFS_unsteady_*.m).

Domain [0, L] x [-1/2, 1/2]; parabolic inflow u = 24 y (1/2 - y) on the upper
half of x = 0 (the lower half is the step face, a wall), 'pres' outflow at
x = L (p_inf = 0), no-slip walls.  Default: 1024 x 256 cells, x uniform,
y clustered towards both walls by a tanh map with stretching factor 1.5 (a
case-defined mesh: mesh_generation.m accepts any x, y from the case),
Re = 800, RK44, direct pressure solve (poisson = 1, as
BFS_unsteady_parameters.m:104 selects; fast diagonalization on the GPU); with
poisson = 3 the PCG settings below apply (relative residual 1e-10).  The
reference's convective
outflow variant is not on the path; the outflow here is the reference's
'pres' condition.
"""
import numpy as np

from ._common import tanh_mesh_1d


def parameters():
    return dict(Re=800.0, Nx=1024, Ny=256, x1=0.0, x2=10.0, y1=-0.5, y2=0.5, sx=1.0, sy=1.0, stretch=1.5,
                order4=0, dt=None, CFL=0.5, t_start=0.0, nsteps=200, RK='RK44',
                poisson=1, CG_acc=1e-10, CG_maxit=50000, p_initial=1, p_add_solve=1)


def BCtype():
    # BFS_unsteady_BCtype.m
    return {'BC_unsteady': 0, 'u': dict(left='dir', right='pres', low='dir', up='dir'),
            'v': dict(left='dir', right='sym', low='dir', up='dir')}


def mesh(options):
    g = options.grid
    x = np.linspace(g.x1, g.x2, g.Nx + 1)
    st = float(options.prm.get('stretch', 0.0))
    y = tanh_mesh_1d(g.y1, g.y2, g.Ny, st) if st > 0 else np.linspace(g.y1, g.y2, g.Ny + 1)
    return x, y


def _inflow(y):
    y = np.asarray(y, float)
    return (y >= 0) * (24.0 * y * (0.5 - y))


def IC(t, options):
    # BFS_unsteady_IC.m: inflow profile everywhere
    g = options.grid
    u = _inflow(g.yu)
    return u, np.zeros((g.Nvx_in, g.Nvy_in)), np.zeros((g.Npx, g.Npy))


def uBC(x, y, t, options):
    # BFS_unsteady_uBC.m
    x = np.atleast_1d(x); y = np.atleast_1d(y)
    if len(x) == 1 and abs(x[0] - options.grid.x1) < 1e-10:
        return _inflow(y)
    return np.zeros(len(x) * len(y))


def vBC(x, y, t, options):
    return np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))


def cfl_dt(options, u, v):
    h = min(np.min(options.grid.hx), np.min(options.grid.hy))
    umax = max(np.max(np.abs(u)), np.max(np.abs(v)), 1.0)
    Re = float(options.prm['Re'])
    return min(float(options.prm.get('CFL', 0.5)) * h / umax, 0.25 * h * h * Re)

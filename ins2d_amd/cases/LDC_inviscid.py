"""Inviscid lid-driven cavity with fourth-order operators
(case_files/LDC_inviscid/LDC_inviscid_*.m): unit square, uniform 20 x 20,
order4 = 1, Re = 1e100, all walls Dirichlet with moving tangential walls
(u = 16 x^2 (1 - x^2) on the lid, u = -1 on the bottom, v = 1 on the left
and v = -1 on the right wall), random initial velocity in [-1, 1].

The reference advances this case with an implicit Gauss-Legendre method
(method 21); the explicit path here takes the case with an explicit tableau
(default RK44).  The reference draws the initial field with MATLAB's rand;
here a seeded generator draws it (``seed``)."""
import numpy as np

from ._common import uniform_mesh


def parameters():
    # LDC_inviscid_parameters.m
    return dict(Re=1e100, Nx=20, Ny=20, x1=0.0, x2=1.0, y1=0.0, y2=1.0, sx=1.0, sy=1.0,
                order4=1, dt=0.01, t_start=0.0, t_end=1.0, RK='RK44',
                poisson=1, p_initial=1, p_add_solve=0, CG_acc=1e-8, CG_maxit=1000, seed=11)


def BCtype():
    # LDC_inviscid_BCtype.m
    return {'BC_unsteady': 0,
            'u': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'},
            'v': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'}}


mesh = uniform_mesh


def IC(t, options):
    # LDC_inviscid_IC.m: u, v uniform in [-1, 1], p = 0
    g = options.grid
    rng = np.random.default_rng(int(options.prm.get('seed', 11)))
    u = -1.0 + 2.0 * rng.random((g.Nux_in, g.Nuy_in))
    v = -1.0 + 2.0 * rng.random((g.Nvx_in, g.Nvy_in))
    return u, v, np.zeros((g.Npx, g.Npy))


def uBC(x, y, t, options):
    # LDC_inviscid_uBC.m
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    g = options.grid
    if len(y) == 1 and abs(y[0] - g.y2) < 1e-10:
        return 16.0 * x ** 2 * (1.0 - x ** 2)
    if len(y) == 1 and abs(y[0] - g.y1) < 1e-10:
        return -np.ones(len(x) * len(y))
    return np.zeros(len(x) * len(y))


def vBC(x, y, t, options):
    # LDC_inviscid_vBC.m
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    g = options.grid
    if len(x) == 1 and abs(x[0] - g.x2) < 1e-10:
        return -np.ones(len(x) * len(y))
    if len(x) == 1 and abs(x[0] - g.x1) < 1e-10:
        return np.ones(len(x) * len(y))
    return np.zeros(len(x) * len(y))

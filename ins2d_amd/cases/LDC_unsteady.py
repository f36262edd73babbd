"""Lid-driven cavity (case_files/LDC_unsteady/LDC_unsteady_*.m).

Defaults are those of case_files/LDC_unsteady/LDC_unsteady_parameters.m
(RK44, direct pressure solve) at 20x20, Re=100, dt=0.01, t=1.  The reference
run logged in results/LDC_unsteady_1.000e+02_20x20_0 is reproduced with
RK='RK33P2' and the direct solve: its log matches that tableau row by row, not
RK44.
"""
import numpy as np

from ._common import uniform_mesh


def parameters():
    return dict(Re=100.0, Nx=20, Ny=20, x1=0.0, x2=1.0, y1=0.0, y2=1.0, sx=1.0, sy=1.0,
                order4=0, dt=0.01, t_start=0.0, t_end=1.0, RK='RK44',
                poisson=1, p_initial=1, p_add_solve=1, CG_acc=1e-8, CG_maxit=1000)


def BCtype():
    # LDC_unsteady_BCtype.m:3-20
    return {'BC_unsteady': 0,
            'u': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'},
            'v': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'}}


mesh = uniform_mesh


def IC(t, options):
    # LDC_unsteady_IC.m:12-17: fluid at rest
    g = options.grid
    return np.zeros((g.Nux_in, g.Nuy_in)), np.zeros((g.Nvx_in, g.Nvy_in)), np.zeros((g.Npx, g.Npy))


def uBC(x, y, t, options):
    # LDC_unsteady_uBC.m:4-11: u = lid speed on y = y2, else 0
    x = np.atleast_1d(x); y = np.atleast_1d(y)
    n = len(x) * len(y)
    if len(y) == 1 and abs(y[0] - options.grid.y2) < 1e-10:
        return np.full(n, float(getattr(options, 'lid_speed', 1.0)))
    return np.zeros(n)


def vBC(x, y, t, options):
    return np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))

"""Made-up fourth-order cases with walls for the tests: every boundary vector
of the fourth-order operators is nonzero in them (the inviscid cavity of the
reference has zero normal wall velocities, so its interpolation and
divergence boundary vectors vanish)."""
from types import SimpleNamespace

import numpy as np

from ins2d_amd.cases._common import tanh_mesh_1d


def wall_case(pery=False, N=(14, 12), Re=200.0):
    """Walls all round (or walls in x, periodic in y) on a tanh-stretched
    grid, with smooth nonzero wall velocities on every wall.  Opposite walls
    carry the same normal velocity profile, so the net flux through the
    boundary vanishes and a divergence-free field exists."""
    side_y = 'per' if pery else 'dir'

    def parameters():
        return dict(Re=Re, Nx=N[0], Ny=N[1], x1=0.0, x2=1.0, y1=0.0, y2=1.0, order4=1, dt=0.002, t_start=0.0,
                    t_end=0.01, RK='RK44', poisson=1, p_initial=1, p_add_solve=1)

    def BCtype():
        return {'BC_unsteady': 0,
                'u': {'left': 'dir', 'right': 'dir', 'low': side_y, 'up': side_y},
                'v': {'left': 'dir', 'right': 'dir', 'low': side_y, 'up': side_y}}

    def mesh(opts):
        g = opts.grid
        x = tanh_mesh_1d(0.0, 1.0, g.Nx, 1.3)
        y = np.linspace(0.0, 1.0, g.Ny + 1) if pery else tanh_mesh_1d(0.0, 1.0, g.Ny, 1.1)
        return x, y

    def IC(t, opts):
        g = opts.grid
        rng = np.random.default_rng(5)
        return (rng.uniform(-1, 1, (g.Nux_in, g.Nuy_in)), rng.uniform(-1, 1, (g.Nvx_in, g.Nvy_in)),
                np.zeros((g.Npx, g.Npy)))

    def uBC(x, y, t, opts):
        x = np.atleast_1d(np.asarray(x, float))
        y = np.atleast_1d(np.asarray(y, float))
        X, Y = np.meshgrid(x, y, indexing='ij')
        return (0.3 + np.sin(2 * np.pi * Y) * (1 + X * (1 - X)) + 0.5 * X * (1 - X)).ravel(order='F')

    def vBC(x, y, t, opts):
        x = np.atleast_1d(np.asarray(x, float))
        y = np.atleast_1d(np.asarray(y, float))
        X, Y = np.meshgrid(x, y, indexing='ij')
        return (-0.2 + 0.7 * np.sin(3 * X) + Y * (1 - Y) * np.cos(2 * np.pi * Y + X)).ravel(order='F')

    return SimpleNamespace(name='walls', parameters=parameters, BCtype=BCtype, mesh=mesh, IC=IC, uBC=uBC, vBC=vBC,
                           dudtBC=None, dvdtBC=None, module=None)

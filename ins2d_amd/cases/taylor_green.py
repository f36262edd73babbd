"""Taylor-Green vortex on the periodic box [0,2pi]^2 with seeded noise.

u = sin x cos y at the u points (x_i, yp_j), v = -cos x sin y at the v points
(xp_i, y_j), plus uniform noise U(-a, a) (a = ``noise``, seed ``seed``).  The
noise breaks the discrete divergence, which the initial projection of
preprocessing/input_check.m:95-111 removes.  The time step is fixed from a CFL
number on the initial field, dt = CFL*h/max|u|, because the reference has no
CFL controller for explicit RK (unsteady/set_timestep.m:5).
"""
import math

import numpy as np

from ._common import periodic, uniform_mesh, zeros_bc


def parameters():
    return dict(Re=2000.0, Nx=1024, Ny=1024, x1=0.0, x2=2 * math.pi, y1=0.0, y2=2 * math.pi, sx=1.0, sy=1.0,
                order4=0, dt=None, CFL=0.5, t_start=0.0, nsteps=2000, RK='RK44',
                poisson=6, p_initial=1, p_add_solve=1, noise=1e-3, seed=1234)


def BCtype():
    return periodic()


mesh = uniform_mesh


def IC(t, options):
    g = options.grid
    prm = options.prm
    rng = np.random.default_rng(int(prm.get('seed', 1234)))
    a = float(prm.get('noise', 0.0))
    u = np.sin(g.xu) * np.cos(g.yu)
    v = -np.cos(g.xv) * np.sin(g.yv)
    if a > 0:
        u = u + rng.uniform(-a, a, size=u.shape)
        v = v + rng.uniform(-a, a, size=v.shape)
    return u, v, np.zeros((g.Npx, g.Npy))


def cfl_dt(options, u, v):
    h = min(np.min(options.grid.hx), np.min(options.grid.hy))
    umax = max(np.max(np.abs(u)), np.max(np.abs(v)))
    return float(options.prm.get('CFL', 0.5)) * h / umax


uBC = zeros_bc
vBC = zeros_bc

"""Taylor-Green vortex with time-dependent Dirichlet boundary values
(case_files/TG_temporal/TG_temporal_*.m).

Domain [0.25, 2.25]^2, Re = 100, uniform 20x20 mesh, Dirichlet on every side
with BC_unsteady = 1.  The boundary values are the exact decaying vortex
u = -sin(pi x) cos(pi y) e^{-2 pi^2 t/Re}, v = cos(pi x) sin(pi y) e^{-2 pi^2 t/Re}
(TG_temporal_uBC.m:7, TG_temporal_vBC.m:6) and their time derivatives
(TG_temporal_dudtBC.m:7, TG_temporal_dvdtBC.m:5).  The reference runs it with
the one-leg method (TG_temporal_parameters.m:84); here the explicit RK path
(method 20, RK44) is the one integrated.
"""
import math

import numpy as np

from ._common import uniform_mesh


def parameters():
    # TG_temporal_parameters.m:9-24,67-70,85,111-122 (j = 1 of the mesh/dt lists)
    return dict(Re=100.0, Nx=20, Ny=20, x1=0.25, x2=2.25, y1=0.25, y2=2.25, sx=1.0, sy=1.0,
                order4=0, dt=0.1, t_start=0.0, t_end=1.0, RK='RK44',
                poisson=1, p_initial=1, p_add_solve=1, CG_acc=1e-8, CG_maxit=1000)


def BCtype():
    # TG_temporal_BCtype.m:3-21
    return {'BC_unsteady': 1,
            'u': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'},
            'v': {'left': 'dir', 'right': 'dir', 'low': 'dir', 'up': 'dir'}}


mesh = uniform_mesh


def IC(t, options):
    # TG_temporal_IC.m:11-16 (the exact field at t = 0)
    g = options.grid
    u = -np.sin(math.pi * g.xu) * np.cos(math.pi * g.yu)
    v = np.cos(math.pi * g.xv) * np.sin(math.pi * g.yv)
    p = 0.25 * (np.cos(2 * math.pi * g.xpp) + np.cos(2 * math.pi * g.ypp))
    return u, v, p


def _outer(f, x, y):
    # the boundary functions are called with one scalar and one vector argument
    return np.multiply.outer(f[0](math.pi * np.atleast_1d(x)), f[1](math.pi * np.atleast_1d(y))).ravel()


def _decay(t, options):
    return math.exp(-2 * math.pi ** 2 * t / options.fluid.Re)


def uBC(x, y, t, options):
    return -_outer((np.sin, np.cos), x, y) * _decay(t, options)


def vBC(x, y, t, options):
    return _outer((np.cos, np.sin), x, y) * _decay(t, options)


def dudtBC(x, y, t, options):
    return (2 * math.pi ** 2 / options.fluid.Re) * _outer((np.sin, np.cos), x, y) * _decay(t, options)


def dvdtBC(x, y, t, options):
    return (-2 * math.pi ** 2 / options.fluid.Re) * _outer((np.cos, np.sin), x, y) * _decay(t, options)


def exact(options, t):
    """The decaying vortex at time t on the u, v and p points (the field the
    boundary functions sample); pressure decays with e^{-4 pi^2 t/Re}."""
    g = options.grid
    e = _decay(t, options)
    u = -np.sin(math.pi * g.xu) * np.cos(math.pi * g.yu) * e
    v = np.cos(math.pi * g.xv) * np.sin(math.pi * g.yv) * e
    p = 0.25 * (np.cos(2 * math.pi * g.xpp) + np.cos(2 * math.pi * g.ypp)) * e * e
    return u, v, p

"""Inviscid/viscous double shear layer on [0,2pi]^2, periodic
(case_files/shear_layer/*.m, testsuite/shear_layer/*.m).

``u_shift`` adds a uniform stream to u (the "+1" runs logged in
results/shear_layer01 and shear_layer03: total u-momentum 4*pi^2).
"""
import math

import numpy as np

from ._common import periodic, uniform_mesh, zeros_bc


def parameters():
    return dict(Re=1e100, Nx=40, Ny=40, x1=0.0, x2=2 * math.pi, y1=0.0, y2=2 * math.pi, sx=1.0, sy=1.0,
                order4=0, dt=0.01, t_start=0.0, t_end=4.0, RK='RK44',
                poisson=6, p_initial=1, p_add_solve=1, CG_acc=1e-8, CG_maxit=1000, u_shift=0.0)


def BCtype():
    return periodic()


mesh = uniform_mesh


def IC(t, options):
    # shear_layer_IC.m:9-14
    g = options.grid
    d = math.pi / 15
    e = 0.05
    yu = g.yu
    u = np.tanh((yu - math.pi / 2) / d) * (yu <= math.pi) + np.tanh((3 * math.pi / 2 - yu) / d) * (yu > math.pi)
    u = u + float(options.prm.get('u_shift', 0.0))
    v = e * np.sin(g.xv)
    return u, v, np.zeros((g.Npx, g.Npy))


uBC = zeros_bc
vBC = zeros_bc

"""Helpers shared by the case plugins."""
import math

import numpy as np


def nonuniform_grid(deltaz, z_low, z_up, sz):
    """Grid lines from z_low to z_up with first spacing deltaz and geometric
    stretch factor sz (libs/nonuniform_grid.m:1-64)."""
    eps = 1e-12
    if sz == 1:
        n = (z_up - z_low) / deltaz
        if abs(n - round(n)) > eps:
            deltaz = (z_up - z_low) / math.floor(n)
        nint = int(math.floor((z_up - z_low) / deltaz + 1e-10))
        return z_low + deltaz * np.arange(nint + 1, dtype=float)
    z = [z_low]
    i = 1
    while z[i - 1] + deltaz * (sz ** i) <= z_up:
        i += 1
        z.append(z[i - 2] + deltaz * (sz ** (i - 2)))
    S = z_up - z[0]
    n = len(z) - 1
    s = [sz, sz ** 2]
    i = 0
    while abs(s[i + 1] - s[i]) > eps:
        fa = S * (s[i + 1] - 1) - deltaz * (s[i + 1] ** n - 1)
        fb = S * (s[i] - 1) - deltaz * (s[i] ** n - 1)
        s.append(s[i + 1] - (s[i + 1] - s[i]) * fa / (fa - fb))
        i += 1
    zz = np.zeros(n + 1)
    zz[0] = z_low
    for k in range(n):
        zz[k + 1] = zz[k] + deltaz * s[-1] ** k
    return zz


def uniform_mesh(options):
    """The ``<case>_mesh.m`` pattern shared by the cavity/periodic cases
    (e.g. case_files/LDC_unsteady/LDC_unsteady_mesh.m)."""
    g = options.grid
    x = nonuniform_grid((g.x2 - g.x1) / g.Nx, g.x1, g.x2, g.sx)
    y = nonuniform_grid((g.y2 - g.y1) / g.Ny, g.y1, g.y2, g.sy)
    return x, y


def tanh_mesh_1d(z1, z2, n, stretch, both_ends=True):
    """Grid lines clustered towards the ends by a tanh map (a case-defined
    mesh: mesh_generation.m:7-8 accepts any x, y from the case)."""
    s = np.linspace(-1.0, 1.0, n + 1) if both_ends else np.linspace(0.0, 1.0, n + 1)
    if both_ends:
        m = np.tanh(stretch * s) / np.tanh(stretch)
        return z1 + (z2 - z1) * (m + 1) / 2
    m = np.tanh(stretch * (s - 1)) / np.tanh(stretch) + 1
    return z1 + (z2 - z1) * m


def periodic(**sides):
    return {'BC_unsteady': 0,
            'u': {'left': 'per', 'right': 'per', 'low': 'per', 'up': 'per'},
            'v': {'left': 'per', 'right': 'per', 'low': 'per', 'up': 'per'}}


def zeros_bc(x, y, t, options):
    return np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))

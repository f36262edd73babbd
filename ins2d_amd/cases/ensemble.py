"""The batched-ensemble workload: cavity and channel instances of uneven
size, aspect ratio, Reynolds number and driving speed (seeded).

Instance k: long side L ~ log-uniform[L_min, L_max] cells, aspect ratio
a ~ U[1, 4], short side round(L / a) (at least 8); domain [0, a] x [0, 1]
with the long side along x.  Half are lid-driven cavities (all walls no-slip,
lid speed U on y = 1; case_files/LDC_unsteady), half are channels (parabolic
inflow 4 U y (1 - y) at x = 0, 'pres' outflow at x = a, no-slip walls; the
BCtype pattern of case_files/BFS_unsteady).  Re ~ U[100, 5000],
U ~ U[0.5, 1.5].  The fixed time step honours both the convective CFL and the
explicit viscous limit: dt = min(CFL h / U, 0.25 h^2 Re).
"""
import math
from types import SimpleNamespace

import numpy as np

from ._common import uniform_mesh

CAVITY = {'BC_unsteady': 0, 'u': dict(left='dir', right='dir', low='dir', up='dir'),
          'v': dict(left='dir', right='dir', low='dir', up='dir')}
CHANNEL = {'BC_unsteady': 0, 'u': dict(left='dir', right='pres', low='dir', up='dir'),
           'v': dict(left='dir', right='sym', low='dir', up='dir')}


def member(kind, Nx, Ny, aspect, Re, U, CFL=0.5, RK='RK44', cg_tol=1e-10, cg_maxit=20000, precond='jacobi'):
    prm = dict(Re=float(Re), Nx=int(Nx), Ny=int(Ny), x1=0.0, x2=float(aspect), y1=0.0, y2=1.0, sx=1.0, sy=1.0,
               order4=0, RK=RK, poisson=3, CG_acc=cg_tol, CG_maxit=cg_maxit, CG_precond=precond, p_initial=1,
               p_add_solve=1, t_start=0.0, nsteps=200, speed=float(U), kind=kind)
    h = min(aspect / Nx, 1.0 / Ny)
    prm['dt'] = min(CFL * h / U, 0.25 * h * h * Re)

    def parameters():
        return dict(prm)

    def BCtype():
        return CAVITY if kind == 'cavity' else CHANNEL

    def IC(t, options):
        g = options.grid
        if kind == 'cavity':
            u = np.zeros((g.Nux_in, g.Nuy_in))
        else:
            u = 4.0 * U * g.yu * (1.0 - g.yu)
        return u, np.zeros((g.Nvx_in, g.Nvy_in)), np.zeros((g.Npx, g.Npy))

    def uBC(x, y, t, options):
        x = np.atleast_1d(x); y = np.atleast_1d(y)
        n = len(x) * len(y)
        if kind == 'cavity':
            if len(y) == 1 and abs(y[0] - 1.0) < 1e-10:
                return np.full(n, U)
            return np.zeros(n)
        if len(x) == 1 and abs(x[0]) < 1e-10:
            return 4.0 * U * y * (1.0 - y)
        return np.zeros(n)

    def vBC(x, y, t, options):
        return np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))

    return SimpleNamespace(name=f'ensemble_{kind}', parameters=parameters, BCtype=BCtype, mesh=uniform_mesh,
                           IC=IC, uBC=uBC, vBC=vBC, module=None)


def members(n=512, seed=4242, L_min=48, L_max=512, **kw):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        L = int(round(math.exp(rng.uniform(math.log(L_min), math.log(L_max)))))
        a = float(rng.uniform(1.0, 4.0))
        short = max(8, int(round(L / a)))
        kind = 'cavity' if k % 2 == 0 else 'channel'
        Re = float(rng.uniform(100.0, 5000.0))
        U = float(rng.uniform(0.5, 1.5))
        out.append(member(kind, L, short, L / short, Re, U, **kw))
    return out

"""Host side of the drop-in: the reference's driver (main.m) and its
``options`` plumbing restated over the libins2d_hip C ABI.

``main(case_name)`` follows main.m:68-235: read the case parameters, build the
mesh from the case's ``mesh`` function, take the boundary-condition types,
build the operators (on the GPU), set the initial condition, sample the
boundary values, run input_check and the explicit RK time loop
(solver_unsteady.m:72-133), and record the per-step conservation log of
process_iteration.m:112-119.  Everything numerical runs in libins2d_hip;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
from types import SimpleNamespace

import numpy as np

from . import _lib
from .cases import load as load_case
from .rk import get_rk_method

BC_KIND = {'dir': 0, 'per': 1, 'pres': 2, 'sym': 3}
POISSON_FFT = 6


class Ins2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'libins2d_hip error {code}: {msg}')
        self.code = code


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check_bc(bct):
    # boundary_conditions.m:17-40
    allowed = {('u', 'left'): ('dir', 'per', 'pres'), ('u', 'right'): ('dir', 'per', 'pres'),
               ('u', 'low'): ('dir', 'per', 'sym'), ('u', 'up'): ('dir', 'per', 'sym'),
               ('v', 'left'): ('dir', 'per', 'sym'), ('v', 'right'): ('dir', 'per', 'sym'),
               ('v', 'low'): ('dir', 'per', 'pres'), ('v', 'up'): ('dir', 'per', 'pres')}
    for (c, side), ok in allowed.items():
        if bct[c][side] not in ok:
            raise ValueError(f'wrong BC for {c}-{side}')


def bc_samples(case, opts, t):
    """Boundary values sampled on the point sets of set_bc_vectors.m:53-71,
    in the order the C ABI takes them."""
    g = opts.grid
    uBC, vBC = case.uBC, case.vBC
    x, y, xin, yin, xp, yp = g.x, g.y, g.xin, g.yin, g.xp, g.yp
    s = [uBC(x, y[:1], t, opts), uBC(x, y[-1:], t, opts),
         uBC(xin, y[:1], t, opts), uBC(xin, y[-1:], t, opts),
         uBC(x[:1], yp, t, opts), uBC(x[-1:], yp, t, opts),
         vBC(x[:1], y, t, opts), vBC(x[-1:], y, t, opts),
         vBC(xp, y[:1], t, opts), vBC(xp, y[-1:], t, opts),
         vBC(x[:1], yin, t, opts), vBC(x[-1:], yin, t, opts)]
    return [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()) for a in s]


def bc_rates(case, opts, t):
    """Time derivatives of the boundary values on the point sets of
    set_bc_vectors.m:73-78 (dudtLe_i, dudtRi_i, dvdtLo_i, dvdtUp_i)."""
    g = opts.grid
    s = [case.dudtBC(g.x[:1], g.yp, t, opts), case.dudtBC(g.x[-1:], g.yp, t, opts),
         case.dvdtBC(g.xp, g.y[:1], t, opts), case.dvdtBC(g.xp, g.y[-1:], t, opts)]
    return [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()) for a in s]


def unsteady_samples(case, opts, c, tn, dt, nsteps):
    """Boundary samples for nsteps RK steps from tn with time-dependent
    boundary values, in the layout of ins2d_step_unsteady: per step the sets
    at tn and at the stage times tn + c_i dt (i = 2..s) and tn + dt
    (time_ERK.m:62-64,106-109), and the rates at tn + dt
    (pressure_additional_solve.m:8-15).  The step start times accumulate as
    solver_unsteady.m does (t = t + dt)."""
    sets, rates = [], []
    t = tn
    for _ in range(nsteps):
        for ti in [t] + [t + ci * dt for ci in c[1:]] + [t + dt]:
            sets.extend(bc_samples(case, opts, ti))
        t = t + dt
        rates.extend(bc_rates(case, opts, t))
    return np.concatenate(sets), np.concatenate(rates)


class Solver:
    """One libins2d_hip handle: operators, state and solver buffers on the GPU."""

    def __init__(self, x, y, bct, Re, RK='RK44', order4=0, poisson=POISSON_FFT, cg_tol=1e-8, cg_maxit=1000,
                 p_add_solve=1, fft_symbol=0, kernels=0, alfa=81.0, beta=9 / 8, cg_precond='jacobi'):
        lib = _lib.load()
        self.lib = lib
        _check_bc(bct)
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        A, b, c = get_rk_method(RK) if isinstance(RK, str) else RK
        self._A = np.ascontiguousarray(A, dtype=np.float64)
        self._b = np.ascontiguousarray(b, dtype=np.float64)
        self._c = np.ascontiguousarray(c, dtype=np.float64)
        d = _lib.CaseDesc()
        d.Nx, d.Ny = len(self.x) - 1, len(self.y) - 1
        d.x, d.y = _dp(self.x), _dp(self.y)
        for k, side in enumerate(('left', 'right', 'low', 'up')):
            d.bc_u[k] = BC_KIND[bct['u'][side]]
            d.bc_v[k] = BC_KIND[bct['v'][side]]
        d.Re = float(Re)
        d.order4 = int(order4)
        d.alfa, d.beta = float(alfa), float(beta)
        d.rk_stages = len(self._b)
        d.rk_A, d.rk_b, d.rk_c = _dp(self._A), _dp(self._b), _dp(self._c)
        d.poisson = int(poisson)
        d.fft_symbol = int(fft_symbol)
        d.cg_tol = float(cg_tol)
        d.cg_maxit = int(cg_maxit)
        d.p_add_solve = int(p_add_solve)
        d.kernels = int(kernels)
        d.cg_precond = {'jacobi': 0, 'ic0': 1, 'two_level': 2, 'fdm': 3}[cg_precond]
        h = ctypes.c_void_p()
        rc = lib.ins2d_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            raise Ins2dError(rc, lib.ins2d_last_error(None).decode())
        self.h = h
        info = np.zeros(7, dtype=np.int32)
        Nx, Ny = d.Nx, d.Ny
        self.xin = np.zeros(Nx + 1)
        self.yin = np.zeros(Ny + 1)
        self._call(lib.ins2d_grid_info, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(self.xin), _dp(self.yin))
        self.Nux_in, self.Nuy_in, self.Nvx_in, self.Nvy_in, self.Npx, self.Npy, self.path = (int(v) for v in info)
        if self.path:   # periodic: interior points are x(1:end-1), y(1:end-1)
            self.xin = self.x[:-1].copy()
            self.yin = self.y[:-1].copy()
        else:
            self.xin = self.xin[:self.Nux_in]
            self.yin = self.yin[:self.Nvy_in]
        self.Nu = self.Nux_in * self.Nuy_in
        self.Nv = self.Nvx_in * self.Nvy_in
        self.NV = self.Nu + self.Nv
        self.Np = self.Npx * self.Npy

    # ------------------------------------------------------------------
    def _call(self, fn, *args):
        rc = fn(self.h, *args)
        if rc != 0:
            raise Ins2dError(rc, self.lib.ins2d_last_error(self.h).decode())

    def close(self):
        if getattr(self, 'h', None):
            self.lib.ins2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bc(self, samples, t=0.0):
        arrs = (ctypes.POINTER(ctypes.c_double) * 12)(*[_dp(s) for s in samples])
        self._keep = samples
        self._call(self.lib.ins2d_set_bc, ctypes.c_double(t), arrs)

    def set_state(self, V=None, p=None):
        V = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        if V is not None and V.size != self.NV:
            raise ValueError('V has the wrong size')
        if p is not None and p.size != self.Np:
            raise ValueError('p has the wrong size')
        self._call(self.lib.ins2d_set_state, None if V is None else _dp(V), None if p is None else _dp(p))

    def get_state(self):
        V = np.zeros(self.NV)
        p = np.zeros(self.Np)
        self._call(self.lib.ins2d_get_state, _dp(V), _dp(p))
        return V, p

    def rhs(self, V, p, t=0.0):
        V = np.ascontiguousarray(V, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        F = np.zeros(self.NV)
        mr = ctypes.c_double()
        self._call(self.lib.ins2d_rhs, _dp(V), _dp(p), ctypes.c_double(t), _dp(F), ctypes.byref(mr))
        return mr.value, F

    def poisson(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        dp = np.zeros(self.Np)
        it = ctypes.c_int()
        rel = ctypes.c_double()
        self._call(self.lib.ins2d_poisson, _dp(f), _dp(dp), ctypes.byref(it), ctypes.byref(rel))
        return dp, it.value, rel.value

    def input_check(self, t=0.0, p_initial=1):
        self._call(self.lib.ins2d_input_check, ctypes.c_double(t), ctypes.c_int(p_initial))

    def step(self, tn, dt, nsteps=1):
        self._call(self.lib.ins2d_step, ctypes.c_double(tn), ctypes.c_double(dt), ctypes.c_int(nsteps))

    def set_bc_rate(self, rates):
        """Time derivatives of the boundary values (bc_rates) at the time of
        the last set_bc; None clears them."""
        if rates is None:
            self._call(self.lib.ins2d_set_bc_rate, None)
            return
        arrs = (ctypes.POINTER(ctypes.c_double) * 4)(*[_dp(r) for r in rates])
        self._keep_rates = rates
        self._call(self.lib.ins2d_set_bc_rate, arrs)

    def step_unsteady(self, tn, dt, nsteps, samples, rates):
        """nsteps with time-dependent boundary values; samples and rates as
        unsteady_samples returns them."""
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        rates = np.ascontiguousarray(rates, dtype=np.float64)
        self._call(self.lib.ins2d_step_unsteady, ctypes.c_double(tn), ctypes.c_double(dt), ctypes.c_int(nsteps),
                   _dp(samples), _dp(rates))

    def diagnostics(self, t=0.0):
        out = np.zeros(5)
        self._call(self.lib.ins2d_diagnostics, ctypes.c_double(t), _dp(out))
        return tuple(float(v) for v in out)

    def solver_stats(self):
        it = ctypes.c_int()
        rel = ctypes.c_double()
        tot = ctypes.c_longlong()
        self._call(self.lib.ins2d_solver_stats, ctypes.byref(it), ctypes.byref(rel), ctypes.byref(tot))
        return it.value, rel.value, tot.value

    def pressure_rhs(self, t=0.0):
        """f = M Om^-1 (-conv + diff) for the resident velocity: the right-hand
        side whose solution is the consistent pressure."""
        f = np.zeros(self.Np)
        self._call(self.lib.ins2d_pressure_rhs, ctypes.c_double(t), _dp(f))
        return f

    def residual(self, f, x):
        """True relative residual ||f - A x|| / ||f||."""
        rel = ctypes.c_double()
        self._call(self.lib.ins2d_residual, _dp(np.ascontiguousarray(f, dtype=np.float64)),
                   _dp(np.ascontiguousarray(x, dtype=np.float64)), ctypes.byref(rel))
        return rel.value

    def set_poisson_accel(self, net=None, warm_start=False, two_level=False, agg=0, coarse_input=False,
                          coarse_start=False, sweeps=0):
        """Accelerated PCG (modes 3/4): learned start vector from ``net``
        (a warmstart.WarmStartNet on this pressure grid) and/or the two-level
        preconditioner; ``coarse_input``: the network also takes the coarse-grid
        solution and predicts its correction (needs both); ``coarse_start``:
        start from the coarse-grid solution alone (two-level, no network);
        ``sweeps``: damped Jacobi sweep pairs after a coarse-input start (0:
        the library's default).  All off restores the plain PCG."""
        flags = ((1 if warm_start else 0) | (2 if two_level else 0) | (4 if coarse_input else 0) |
                 (8 if coarse_start else 0) | (int(sweeps) << 8))
        self._accel_net = net if warm_start else None   # keep the network alive
        self._call(self.lib.ins2d_set_poisson_accel, net._h if warm_start else None, ctypes.c_int(flags),
                   ctypes.c_int(agg))

    def pressure_rhs_dev(self, f_ptr, t=0.0):
        """The pressure right-hand side of the resident velocity into the
        device array at f_ptr."""
        self._call(self.lib.ins2d_pressure_rhs_dev, ctypes.c_double(t), ctypes.c_void_p(f_ptr))

    def poisson_dev(self, f_ptr, dp_ptr):
        """A dp = f on device arrays."""
        self._call(self.lib.ins2d_poisson_dev, ctypes.c_void_p(f_ptr), ctypes.c_void_p(dp_ptr))

    def poisson_coarse_dev(self, f_ptr, xc_ptr):
        """The coarse-grid solution of f into the device array at xc_ptr."""
        self._call(self.lib.ins2d_poisson_coarse_dev, ctypes.c_void_p(f_ptr), ctypes.c_void_p(xc_ptr))

    def poisson_coarse(self, f):
        """The coarse-grid solution P Ac^-1 P^T f of the two-level
        preconditioner (set_poisson_accel(two_level=True) first)."""
        xc = np.zeros(self.Np)
        self._call(self.lib.ins2d_poisson_coarse, _dp(np.ascontiguousarray(f, dtype=np.float64)), _dp(xc))
        return xc

    def accel_stats(self):
        fb = ctypes.c_longlong()
        agg = ctypes.c_int()
        self._call(self.lib.ins2d_poisson_accel_stats, ctypes.byref(fb), ctypes.byref(agg))
        return fb.value, agg.value

    def profile(self, enable):
        self._call(self.lib.ins2d_profile, ctypes.c_int(int(enable)))

    def profile_read(self, kind):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._call(self.lib.ins2d_profile_read, ctypes.c_int(kind), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def sync(self):
        self._call(self.lib.ins2d_sync)


def _options(case, prm):
    g = SimpleNamespace(x1=prm.get('x1', 0.0), x2=prm.get('x2', 1.0), y1=prm.get('y1', 0.0), y2=prm.get('y2', 1.0),
                        Nx=prm['Nx'], Ny=prm['Ny'], sx=prm.get('sx', 1.0), sy=prm.get('sy', 1.0))
    return SimpleNamespace(grid=g, prm=prm, case_obj=case, fluid=SimpleNamespace(Re=float(prm['Re'])))


def _prepare(case):
    """Parameters, options and mesh of a case (main.m:68-142)."""
    prm = dict(case.parameters())
    opts = _options(case, prm)
    x, y = case.mesh(opts)
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    g = opts.grid
    g.x, g.y = x, y
    g.xp = (x[:-1] + x[1:]) / 2
    g.yp = (y[:-1] + y[1:]) / 2
    g.hx, g.hy = np.diff(x), np.diff(y)
    g.Nx, g.Ny = len(g.hx), len(g.hy)
    bct = case.BCtype()
    opts.unsteady = bool(int(bct.get('BC_unsteady', 0)))
    if opts.unsteady and (case.dudtBC is None or case.dvdtBC is None):
        raise ValueError('BC_unsteady = 1 needs dudtBC and dvdtBC (set_bc_vectors.m:73-78)')
    _check_bc(bct)
    return prm, opts, bct


def _finish(case, opts, sizes, xin, yin):
    """Grid metrics, initial condition, time step and boundary samples
    (operator_mesh.m:401-414, initialize.m, set_bc_vectors.m:53-71)."""
    g = opts.grid
    prm = opts.prm
    g.xin, g.yin = xin, yin
    g.Nux_in, g.Nuy_in, g.Nvx_in, g.Nvy_in, g.Npx, g.Npy = sizes[:6]
    g.xu = np.outer(g.xin, np.ones(g.Nuy_in)); g.yu = np.outer(np.ones(g.Nux_in), g.yp)
    g.xv = np.outer(g.xp, np.ones(g.Nvy_in)); g.yv = np.outer(np.ones(g.Nvx_in), g.yin)
    g.xpp = np.outer(g.xp, np.ones(g.Ny)); g.ypp = np.outer(np.ones(g.Nx), g.yp)
    t = float(prm.get('t_start', 0.0))
    u0, v0, p0 = case.IC(t, opts)
    V0 = np.r_[np.asarray(u0, float).ravel(order='F'), np.asarray(v0, float).ravel(order='F')]
    p0 = np.asarray(p0, float).ravel(order='F')
    dt = prm.get('dt')
    if dt is None:
        dt = case.module.cfl_dt(opts, u0, v0)
    opts.dt = float(dt)
    t_end = prm.get('t_end')
    opts.nt = int(prm['nsteps']) if t_end is None else int(math.ceil((t_end - t) / opts.dt - 1e-12))
    opts.t_start = t
    return V0, p0, bc_samples(case, opts, t)


class Batch:
    """Many cases advanced together by the batched ensemble of libins2d_hip
    (second order, PCG pressure, one shared tableau)."""

    def __init__(self, cases):
        lib = _lib.load()
        self.lib = lib
        self.cases = cases
        self.opts, descs, keep = [], [], []
        RK = None
        for c in cases:
            prm, opts, bct = _prepare(c)
            RK = RK or prm.get('RK', 'RK44')
            if prm.get('RK', 'RK44') != RK:
                raise ValueError('all ensemble members share one RK method')
            A, b, cc = get_rk_method(RK)
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (opts.grid.x, opts.grid.y, A, b, cc)]
            keep.append(arrs)
            d = _lib.CaseDesc()
            d.Nx, d.Ny = opts.grid.Nx, opts.grid.Ny
            d.x, d.y = _dp(arrs[0]), _dp(arrs[1])
            for k, side in enumerate(('left', 'right', 'low', 'up')):
                d.bc_u[k] = BC_KIND[bct['u'][side]]
                d.bc_v[k] = BC_KIND[bct['v'][side]]
            d.Re = float(prm['Re'])
            d.rk_stages = len(b)
            d.rk_A, d.rk_b, d.rk_c = _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4])
            d.poisson = int(prm.get('poisson', 3))
            d.cg_tol = float(prm.get('CG_acc', 1e-10))
            d.cg_maxit = int(prm.get('CG_maxit', 1000))
            d.p_add_solve = int(prm.get('p_add_solve', 1))
            d.cg_precond = {'jacobi': 0, 'ic0': 1, 'two_level': 2, 'fdm': 3}[prm.get('CG_precond', 'jacobi')]
            descs.append(d)
            self.opts.append(opts)
        self._keep = keep
        arr = (_lib.CaseDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        rc = lib.ins2d_batch_create(arr, len(descs), ctypes.byref(h))
        if rc != 0:
            raise Ins2dError(rc, lib.ins2d_batch_last_error(None).decode())
        self.h = h
        self.n = len(cases)
        self.V0, self.p0, self.sizes = [], [], []
        self._samples = []
        for k, (c, opts) in enumerate(zip(cases, self.opts)):
            sz = np.zeros(6, dtype=np.int32)
            xin = np.zeros(opts.grid.Nx + 1)
            yin = np.zeros(opts.grid.Ny + 1)
            self._call(lib.ins2d_batch_info, ctypes.c_int(k), sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       _dp(xin), _dp(yin))
            sizes = [int(v) for v in sz]
            V0, p0, samples = _finish(c, opts, sizes, xin[:sizes[0]], yin[:sizes[3]])
            self.sizes.append(sizes)
            self.V0.append(V0)
            self.p0.append(p0)
            arrs = (ctypes.POINTER(ctypes.c_double) * 12)(*[_dp(s) for s in samples])
            self._samples.append(samples)
            self._call(lib.ins2d_batch_set_bc, ctypes.c_int(k), ctypes.c_double(opts.t_start), arrs)
            self._call(lib.ins2d_batch_set_state, ctypes.c_int(k), _dp(V0), _dp(p0))
        self.dt = np.array([o.dt for o in self.opts], dtype=np.float64)
        self.cells = sum(s[4] * s[5] for s in self.sizes)

    def _call(self, fn, *args):
        rc = fn(self.h, *args)
        if rc != 0:
            raise Ins2dError(rc, self.lib.ins2d_batch_last_error(self.h).decode())

    def input_check(self, p_initial=1):
        self._call(self.lib.ins2d_batch_input_check, ctypes.c_int(p_initial))

    def step(self, nsteps=1):
        self._call(self.lib.ins2d_batch_step, _dp(self.dt), ctypes.c_int(nsteps))

    def get_state(self, k):
        s = self.sizes[k]
        V = np.zeros(s[0] * s[1] + s[2] * s[3])
        p = np.zeros(s[4] * s[5])
        self._call(self.lib.ins2d_batch_get_state, ctypes.c_int(k), _dp(V), _dp(p))
        return V, p

    def autotune(self):
        """Time the stage kernel for a few workgroup shapes; keep the fastest."""
        tx, ty = ctypes.c_int(), ctypes.c_int()
        self._call(self.lib.ins2d_batch_autotune, ctypes.byref(tx), ctypes.byref(ty))
        return tx.value, ty.value

    def set_tile(self, tx, ty):
        """Set the stage kernel's workgroup shape (tx * ty = 256)."""
        self._call(self.lib.ins2d_batch_set_tile, ctypes.c_int(tx), ctypes.c_int(ty))

    def stats(self):
        tot = ctypes.c_longlong()
        mx = ctypes.c_int()
        cells = ctypes.c_longlong()
        self._call(self.lib.ins2d_batch_stats, ctypes.byref(tot), ctypes.byref(mx), ctypes.byref(cells))
        return tot.value, mx.value, cells.value

    def close(self):
        if getattr(self, 'h', None):
            self.lib.ins2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def setup(case, **overrides):
    """main.m:68-186 up to the start of the solver: returns (solver, options,
    V_start, p_start) with the operators built and the boundary samples set."""
    if isinstance(case, str):
        case = load_case(case, **overrides)
    prm, opts, bct = _prepare(case)
    g = opts.grid
    solver = Solver(g.x, g.y, bct, prm['Re'], RK=prm.get('RK', 'RK44'), order4=prm.get('order4', 0),
                    poisson=prm.get('poisson', POISSON_FFT), cg_tol=prm.get('CG_acc', 1e-8),
                    cg_maxit=prm.get('CG_maxit', 1000), p_add_solve=prm.get('p_add_solve', 1),
                    fft_symbol=prm.get('fft_symbol', 0), kernels=prm.get('kernels', 0),
                    alfa=prm.get('alfa', 81.0), beta=prm.get('beta', 9 / 8),
                    cg_precond=prm.get('CG_precond', 'jacobi'))
    sizes = [solver.Nux_in, solver.Nuy_in, solver.Nvx_in, solver.Nvy_in, solver.Npx, solver.Npy]
    V0, p0, samples = _finish(case, opts, sizes, solver.xin, solver.yin)
    solver.set_bc(samples, opts.t_start)
    if opts.unsteady:
        solver.set_bc_rate(bc_rates(case, opts, opts.t_start))
    opts.case_obj = case
    return solver, opts, V0, p0


def main(case_name, nsteps=None, diagnostics=True, path_results=None, restart_n=0, snapshots=False,
         restart_from=None, **overrides):
    """main.m + solver_unsteady.m: returns (V, p, rows) where rows are the
    convergence.txt records (n, dt, t, maxres, maxdiv, umom, vmom, k).

    Output as the reference writes it (options.output, restart.*), when
    path_results is given: convergence.txt (create_files.m:54-56,
    solver_unsteady.m:4-8, process_iteration.m:112-119), a restart file every
    restart_n steps (process_iteration.m:56-57, write_restart.m) and, with
    snapshots, snapshots/Vp_n=<n>.mat every step (process_iteration.m:78-86,
    save_eco = 1).  restart_from: a restart file to continue from (its V, p,
    n, t and dt replace the initial condition, as main.m with restart.load =
    1; input_check is not repeated)."""
    import os
    solver, opts, V0, p0 = setup(case_name, **overrides)
    prm = opts.prm
    t = opts.t_start
    dt = opts.dt
    n0 = 1
    if restart_from is not None:
        r = read_restart(restart_from)
        V0 = np.concatenate([r['uh'], r['vh']])
        p0 = r['p']
        n0, t, dt = int(r['n']), float(r['t']), float(r['dt'])
        if opts.unsteady:
            solver.set_bc(bc_samples(opts.case_obj, opts, t), t)
            solver.set_bc_rate(bc_rates(opts.case_obj, opts, t))
    solver.set_state(V0, p0)
    if restart_from is None:
        solver.input_check(t, int(prm.get('p_initial', 1)))
    rows = [(n0, dt, t) + solver.diagnostics(t)] if diagnostics else []
    nt = opts.nt if nsteps is None else nsteps
    name = case_name if isinstance(case_name, str) else getattr(case_name, 'name', 'case')
    nu = solver.Nu
    for n in range(n0 + 1, n0 + nt + 1):
        if opts.unsteady:
            solver.step_unsteady(t, dt, 1, *unsteady_samples(opts.case_obj, opts, solver._c, t, dt, 1))
        else:
            solver.step(t, dt, 1)
        t = t + dt
        if diagnostics:
            rows.append((n, dt, t) + solver.diagnostics(t))
        if path_results and (restart_n or snapshots):
            V, p = solver.get_state()
            if restart_n and n % restart_n == 0:
                write_restart(path_results, n, t, dt, V, p, nu)
            if snapshots:
                write_snapshot(path_results, n, V, p)
    V, p = solver.get_state()
    solver.close()
    if path_results and diagnostics:
        os.makedirs(path_results, exist_ok=True)
        g = opts.grid
        write_convergence(os.path.join(path_results, 'convergence.txt'), name, g.Nx, g.Ny, float(prm['Re']), rows)
    return V, p, rows


def write_convergence(path, case_name, Nx, Ny, Re, rows):
    """convergence.txt in the layout of create_files.m:54-56 (case, Nx, Ny,
    num2str(Re, '%10.3e')), the column header and first record of
    solver_unsteady.m:4-8 (no trailing blank) and the records of
    process_iteration.m:112-119 (a trailing blank)."""
    with open(path, 'w') as fh:
        # num2str(Re, '%10.3e') drops the padding blanks of the format
        fh.write(f'{case_name} {Nx} {Ny} {("%10.3e" % Re).strip()}\n')
        fh.write('residual and conservation information\n')
        fh.write('n            dt               t                res              maxdiv           umom             vmom             k\n')
        for q, r in enumerate(rows):
            line = '%-10i %16.8e %16.8e %16.8e %16.8e %16.8e %16.8e %16.8e' % tuple(r)
            fh.write(line + ('\n' if q == 0 else ' \n'))


def _numbered(n, filelen=8):
    # process_iteration.m:103-104: n left-padded with zeros to filelen digits
    s = str(int(n))
    return '0' * max(0, filelen - len(s)) + s


def write_restart(path_results, n, t, dt, V, p, nu, filelen=8):
    """restart_<n>.mat as write_restart.m: uh, vh, p (columns), n, t, dt
    (MAT-file version 5, readable by MATLAB's load)."""
    import os
    from scipy.io import savemat
    os.makedirs(path_results, exist_ok=True)
    path = os.path.join(path_results, f'restart_{_numbered(n, filelen)}.mat')
    savemat(path, {'uh': np.asarray(V[:nu], dtype=np.float64), 'vh': np.asarray(V[nu:], dtype=np.float64),
                   'p': np.asarray(p, dtype=np.float64), 'n': float(n), 't': float(t), 'dt': float(dt)},
            oned_as='column')
    return path


def read_restart(path):
    """The variables of a restart file (write_restart), flattened."""
    from scipy.io import loadmat
    m = loadmat(path)
    out = {k: np.asarray(m[k], dtype=np.float64).ravel() for k in ('uh', 'vh', 'p')}
    for k in ('n', 't', 'dt'):
        out[k] = float(np.asarray(m[k]).ravel()[0])
    return out


def write_snapshot(path_results, n, V, p):
    """snapshots/Vp_n=<n>.mat with V and p (process_iteration.m:84-86)."""
    import os
    from scipy.io import savemat
    d = os.path.join(path_results, 'snapshots')
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f'Vp_n={int(n)}.mat')
    savemat(path, {'V': np.asarray(V, dtype=np.float64), 'p': np.asarray(p, dtype=np.float64)}, oned_as='column')
    return path

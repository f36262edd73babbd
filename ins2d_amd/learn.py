"""Training data, trainer and evaluation of the learned Poisson warm start
(no reference counterpart).

* ``generate_pairs``: (f, p) pairs from plain-solver runs of the
  backward-facing step: the velocity is advanced with the direct pressure
  solve, and every ``every`` steps the pressure right-hand side of the
  resident velocity, f = M Om^-1 (-conv + diff), and its exact solution
  p = A^-1 f are recorded (host arrays; for evaluation sets).
* ``generate_pairs_dev``: the same written straight into the network's
  device-resident training set, with the coarse-grid solution
  xc = P Ac^-1 P^T f of the two-level preconditioner as the third array
  (tens of thousands of pairs at 1024 x 256 exceed host memory, not HBM).
* ``train``: Adam on the library's own BF16 forward/backward kernels over the
  device-resident pairs, with linear-combination augmentation (A and the
  coarse solve are linear, so a f1 + b f2 -> a p1 + b p2, a xc1 + b xc2 is
  again an exact pair).
* ``evaluate``: per-solve GPU time (HIP events around the whole solve,
  network inference included) and iterations of the plain Jacobi-PCG against
  the accelerated variants, all iterating to the same relative residual.
"""
import time

import numpy as np

from . import host
from .warmstart import WarmStartNet


def generate_pairs(Nx, Ny, Re_values, steps, every, skip=0, progress=None):
    """List of (f, p, Re, step) from BFS runs at each Re (direct solve)."""
    out = []
    for Re in Re_values:
        solver, opts, V0, p0 = host.setup('BFS_unsteady', Nx=Nx, Ny=Ny, Re=float(Re), poisson=1)
        solver.set_state(V0, p0)
        solver.input_check(0.0, 1)
        t, dt = 0.0, opts.dt
        if skip:
            solver.step(t, dt, skip)
            t += skip * dt
        for n in range(steps // every):
            solver.step(t, dt, every)
            t += every * dt
            f = solver.pressure_rhs(t)
            p, _, _ = solver.poisson(f)
            out.append((f, p, float(Re), skip + (n + 1) * every))
        solver.close()
        if progress:
            progress(f'pairs: Re={Re:.0f} done, {len(out)} total')
    return out


def coarse_solver(Nx, Ny, agg=0):
    """A handle on the BFS pressure grid with the two-level preconditioner,
    used for the coarse-grid solutions P Ac^-1 P^T f."""
    s, _, _, _ = host.setup('BFS_unsteady', Nx=Nx, Ny=Ny, poisson=3, CG_acc=1e-10, CG_maxit=100000)
    s.set_poisson_accel(None, two_level=True, agg=agg)
    return s


def generate_pairs_dev(net, Nx, Ny, Re_values, per_run, every, skip=0, first=0, agg=0, progress=None):
    """Pairs first .. first + len(Re_values) * per_run of the network's
    training set (reserved by the caller): BFS runs at each Re with the direct
    solve; every ``every`` steps f, p = A^-1 f and xc are written into the
    pair's device arrays.  Returns the number of pairs written."""
    cs = coarse_solver(Nx, Ny, agg)
    k = first
    for Re in Re_values:
        solver, opts, V0, p0 = host.setup('BFS_unsteady', Nx=Nx, Ny=Ny, Re=float(Re), poisson=1)
        solver.set_state(V0, p0)
        solver.input_check(0.0, 1)
        t, dt = 0.0, opts.dt
        if skip:
            solver.step(t, dt, skip)
            t += skip * dt
        for _ in range(per_run):
            solver.step(t, dt, every)
            t += every * dt
            f, p, xc = net.data_ptr(k, 0), net.data_ptr(k, 1), net.data_ptr(k, 2)
            solver.pressure_rhs_dev(f, t)
            solver.poisson_dev(f, p)
            cs.poisson_coarse_dev(f, xc)
            k += 1
        solver.close()
        if progress:
            progress(f'device pairs: Re={Re:.0f} done, {k - first} total')
    cs.close()
    return k - first


def rms(a):
    return float(np.sqrt(np.mean(np.square(a))))


def out_scale_from_net(net, n, samples=64, seed=0):
    """Median of rms(p - xc) / rms(f) over a sample of the stored pairs."""
    rng = np.random.default_rng(seed)
    r = []
    for k in rng.choice(n, size=min(samples, n), replace=False):
        f, p, xc = net.data_get(int(k))
        r.append(rms(p - xc) / max(rms(f), 1e-300))
    return float(np.median(r))


def train(net, pairs, steps, batch=4, lr=2e-3, lr_end=1e-4, seed=0, combos=True, progress=None, every=50):
    """Adam steps on host pairs (uploaded once to device memory).  With
    ``combos``, half of each batch are random linear combinations of two
    pairs.  Returns the loss history."""
    n = len(pairs)
    ncombo = batch // 2 if combos else 0
    net.data_reserve(n + ncombo)
    for k, (f, p, *rest) in enumerate(pairs):
        net.data_set(k, f, p, rest[0] if rest and isinstance(rest[0], np.ndarray) else None)
    return train_stored(net, n, steps, batch, lr, lr_end, seed, combos, progress, every)


def train_stored(net, n, steps, batch=4, lr=2e-3, lr_end=1e-4, seed=0, combos=True, progress=None, every=50,
                 relative=False):
    """Adam steps on the first n stored pairs; combinations are formed on the
    device in slots n .. n + batch/2 (reserved by the caller).  Cosine
    learning-rate decay from lr to lr_end.  Returns the loss history."""
    rng = np.random.default_rng(seed)
    ncombo = batch // 2 if combos else 0
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        lr_s = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
        idx = list(rng.integers(0, n, batch - ncombo))
        for c in range(ncombo):
            i, j = rng.choice(n, size=2, replace=False)
            a, b = rng.standard_normal(2)
            net.data_combine(n + c, int(i), int(j), float(a), float(b))
            idx.append(n + c)
        losses.append(net.train(idx, lr=float(lr_s), update=True, relative=relative))
        if progress and (s % every == 0 or s == steps - 1):
            progress(f'train step {s}/{steps}: loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)')
    return losses


def out_scale_of(pairs):
    return float(np.median([rms(p) / max(rms(f), 1e-300) for f, p, *_ in pairs]))


def evaluate(solver, net, rhs_list, variants, progress=None):
    """For each variant (name, options of Solver.set_poisson_accel) and each
    f: GPU milliseconds of the solve (HIP events), iterations, true relative
    residual.  Returns {name: dict(ms=[...], it=[...], rel=[...], fallbacks)}."""
    res = {}
    for name, opt in variants:
        accel = any(opt.get(k) for k in ('warm_start', 'two_level', 'coarse_start'))
        solver.set_poisson_accel(net if opt.get('warm_start') else None, **opt)
        fb0 = solver.accel_stats()[0] if accel else 0
        solver.poisson(rhs_list[0])   # warm-up (first-call set-up, caches)
        ms, its, rels = [], [], []
        for f in rhs_list:
            solver.profile(True)
            x, it, _ = solver.poisson(f)
            solver.sync()
            m, cnt = solver.profile_read(2)
            solver.profile(False)
            ms.append(m)
            its.append(it)
            rels.append(solver.residual(f, x))
        fb = (solver.accel_stats()[0] - fb0) if accel else 0
        res[name] = dict(ms=ms, it=its, rel=rels, fallbacks=fb)
        if progress:
            progress(f'{name}: median {np.median(ms):.3f} ms, {np.median(its):.0f} iterations, '
                     f'max true relres {np.max(rels):.2e}, fallbacks {fb}')
    solver.set_poisson_accel(None)
    return res


def make_net(Nx, Ny, out_scale=1.0, seed=0):
    """A network on the Ny x Nx pressure grid with He-normal weights."""
    net = WarmStartNet(Ny, Nx)
    net.init(seed, out_scale)
    return net

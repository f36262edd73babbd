"""Training data, trainer and evaluation of the learned Poisson warm start
(no reference counterpart).

* ``generate_pairs``: (f, p) pairs from plain-solver runs of the
  backward-facing step: the velocity is advanced with the direct pressure
  solve, and every ``every`` steps the pressure right-hand side of the
  resident velocity, f = M Om^-1 (-conv + diff), and its exact solution
  p = A^-1 f are recorded.
* ``train``: Adam on the library's own BF16 forward/backward kernels, on pairs
  held in device memory, with linear-combination augmentation (A is linear,
  so a f1 + b f2 -> a p1 + b p2 is again an exact pair).
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


def rms(a):
    return float(np.sqrt(np.mean(np.square(a))))


def train(net, pairs, steps, batch=4, lr=2e-3, lr_end=1e-4, seed=0, combos=True, progress=None, every=50):
    """Adam steps on the pairs (uploaded once to device memory).  With
    ``combos``, half of each batch are random linear combinations of two
    pairs.  Returns the loss history."""
    rng = np.random.default_rng(seed)
    n = len(pairs)
    ncombo = batch // 2 if combos else 0
    net.data_reserve(n + ncombo)
    for k, (f, p, *_) in enumerate(pairs):
        net.data_set(k, f, p)
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        lr_s = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
        idx = list(rng.integers(0, n, batch - ncombo))
        for c in range(ncombo):
            i, j = rng.integers(0, n, 2)
            a, b = rng.standard_normal(2)
            fi, pi = pairs[i][0], pairs[i][1]
            fj, pj = pairs[j][0], pairs[j][1]
            # weight by 1/rms so both members contribute
            a, b = a / max(rms(fi), 1e-300), b / max(rms(fj), 1e-300)
            net.data_set(n + c, a * fi + b * fj, a * pi + b * pj)
            idx.append(n + c)
        losses.append(net.train(idx, lr=float(lr_s), update=True))
        if progress and (s % every == 0 or s == steps - 1):
            progress(f'train step {s}/{steps}: loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)')
    return losses


def out_scale_of(pairs):
    return float(np.median([rms(p) / max(rms(f), 1e-300) for f, p, *_ in pairs]))


def evaluate(solver, net, rhs_list, variants, progress=None):
    """For each variant (name, warm_start, two_level) and each f: GPU
    milliseconds of the solve (HIP events), iterations, true relative
    residual.  Returns {name: dict(ms=[...], it=[...], rel=[...], fallbacks)}."""
    res = {}
    for name, ws, tl in variants:
        solver.set_poisson_accel(net if ws else None, warm_start=ws, two_level=tl)
        fb0 = solver.accel_stats()[0] if (ws or tl) else 0
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
        fb = (solver.accel_stats()[0] - fb0) if (ws or tl) else 0
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

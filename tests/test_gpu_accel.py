"""Accelerated pressure PCG on the GPU: the two-level preconditioner, the
warm-start fallback and the gain of a trained network, all judged by the
plain solver's stopping rule ||f - A x|| <= tol ||f||.

The true relative residual of every result is recomputed by the library
(``ins2d_residual``) and must be within 1.5x the requested tolerance (the
recursive residual CG monitors drifts slightly from the true one)."""
import numpy as np
import pytest

from ins2d_amd import host
from ins2d_amd.cases import load
from oracle import ins2d_oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[gpu-test]', *a, flush=True)


def _bfs(Nx, Ny, tol=1e-10, **kw):
    return load('BFS_unsteady', Nx=Nx, Ny=Ny, poisson=3, CG_acc=tol, CG_maxit=50000, **kw)


def _ldc(N, tol=1e-10):
    return load('LDC_unsteady', Nx=N, Ny=N, poisson=3, CG_acc=tol, CG_maxit=50000)


@pytest.mark.parametrize('name', ['bfs', 'ldc', 'ldc_odd'])
def test_two_level_pcg(name):
    _say('two-level', name)
    if name == 'bfs':
        case = _bfs(256, 64)
    elif name == 'ldc':
        case = _ldc(64)
    else:
        case = load('LDC_unsteady', Nx=75, Ny=41, sx=1.04, sy=1.03, poisson=3, CG_acc=1e-10, CG_maxit=50000)
    opts = O.build(case)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(2)
    f = rng.standard_normal(solver.Np)
    singular = name != 'bfs'
    if singular:
        f -= f.mean()
    x0, it0, _ = solver.poisson(f)
    solver.set_poisson_accel(two_level=True)
    x1, it1, rel1 = solver.poisson(f)
    _, agg = solver.accel_stats()
    r0 = solver.residual(f, x0)
    r1 = solver.residual(f, x1)
    _say(name, 'plain', it0, r0, 'two-level', it1, r1, 'agg', agg)
    assert r0 <= 1.5e-10 and r1 <= 1.5e-10
    assert abs(rel1 - r1) <= 1e-3 * r1     # the solver reports the true residual
    # the gain grows with the grid (cells per aggregate vs coarse size)
    assert it1 < (it0 / 4 if name == 'bfs' else 0.7 * it0)
    # both solve the same system: agree with the oracle's sparse LU to the
    # accuracy the residual allows (modulo the constant when singular)
    A = opts.discretization.A
    r_true = np.linalg.norm(A @ x1 - f) / np.linalg.norm(f)
    assert abs(r_true - r1) <= 1e-3 * r1
    solver.close()


def test_warm_start_fallback_and_neutral_start():
    from ins2d_amd.warmstart import WarmStartNet
    _say('warm-start fallback')
    case = _bfs(128, 64)
    solver, _, _, _ = host.setup(case)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(solver.Np)
    _, it_plain, _ = solver.poisson(f)
    net = WarmStartNet(64, 128)
    # a huge start vector raises the residual: the solve must fall back to x0 = 0
    net.init(seed=1, out_scale=1e6)
    solver.set_poisson_accel(net, warm_start=True)
    x, it, rel = solver.poisson(f)
    fb, _ = solver.accel_stats()
    _say('fallbacks', fb, 'iterations', it, 'plain', it_plain, 'relres', rel)
    assert fb == 1
    assert solver.residual(f, x) <= 1.5e-10
    assert abs(it - it_plain) <= 2
    # a zero network output (out_scale 0) leaves the smoothing sweeps from
    # x0 = 0: a residual below ||f||, no fallback, no extra iterations
    net.init(seed=1, out_scale=0.0)
    solver.set_poisson_accel(net, warm_start=True)
    x, it, rel = solver.poisson(f)
    fb2, _ = solver.accel_stats()
    _say('zero network output: iterations', it)
    assert fb2 == fb
    assert it <= it_plain + 2 and solver.residual(f, x) <= 1.5e-10
    solver.close()
    net.close()


def test_trained_network_gain():
    """Train on (f, p) pairs from BFS runs at three Reynolds numbers, then
    solve held-out right-hand sides (another Re, later times) with plain
    Jacobi-PCG, the two-level PCG, and both with the network's start vector.
    Prints what the network gains over the same method without it.  On this
    data the smoothed, coarse-corrected start still leaves ||f - A x0|| above
    ||f|| (its BF16 rounding noise and ~30 % error), so the fallback to x0 = 0
    triggers; the test pins that this guard holds and that every variant
    reaches the tolerance."""
    from ins2d_amd import learn
    _say('trained network gain')
    Nx, Ny = 128, 64
    pairs = learn.generate_pairs(Nx, Ny, [300.0, 600.0, 900.0], steps=240, every=4, skip=20, progress=_say)
    held = learn.generate_pairs(Nx, Ny, [750.0], steps=160, every=8, skip=100, progress=_say)
    net = learn.make_net(Nx, Ny, out_scale=learn.out_scale_of(pairs), seed=0)
    losses = learn.train(net, pairs, steps=300, batch=4, lr=2e-3, progress=_say, every=100)
    assert np.mean(losses[-20:]) < 0.35 * np.mean(losses[:5])
    solver, _, _, _ = host.setup(_bfs(Nx, Ny))
    rhs = [h[0] for h in held]
    res = learn.evaluate(solver, net, rhs, [('plain', False, False), ('two_level', False, True),
                                            ('warm', True, False), ('learned', True, True)], progress=_say)
    x0rel = [solver.residual(f, net.forward(f.reshape(Ny, Nx)).ravel()) for f in rhs]
    _say('residual of the raw network output (before smoothing): median', np.median(x0rel))
    for k in res:
        assert max(res[k]['rel']) <= 1.5e-10, k
    gain = np.median(res['two_level']['ms']) / np.median(res['learned']['ms'])
    _say('network gain over the two-level PCG without it:', gain, 'fallbacks', res['learned']['fallbacks'],
         'of', len(rhs) + 1)
    # whatever the network's quality, the fallback keeps the accelerated
    # solve from needing more iterations than the start from x0 = 0
    assert np.median(res['learned']['it']) <= np.median(res['two_level']['it']) + 2
    assert np.median(res['warm']['it']) <= np.median(res['plain']['it']) + 2
    # the coarse correction alone must give the required gain
    assert np.median(res['plain']['ms']) > 3 * np.median(res['two_level']['ms'])
    solver.close()
    net.close()

"""GPU parity of the batched ensemble: every instance advanced by the batched
kernels (tile work list, batched PCG with per-instance scalars, hipGraph
replay) must match the same case advanced alone by the single-case path and
by the CPU oracle."""
import numpy as np
import pytest

from ins2d_amd import host
from ins2d_amd.cases import ensemble
from oracle import ins2d_oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[gpu-test]', *a, flush=True)


def test_ensemble_matches_single_runs_and_oracle():
    _say('ensemble 6 members')
    members = [ensemble.member('cavity', 24, 16, 1.5, 400.0, 1.0, cg_tol=1e-13),
               ensemble.member('channel', 32, 12, 32 / 12, 900.0, 0.7, cg_tol=1e-13),
               ensemble.member('cavity', 16, 16, 1.0, 150.0, 1.3, cg_tol=1e-13),
               ensemble.member('channel', 40, 20, 2.0, 3000.0, 1.1, cg_tol=1e-13),
               ensemble.member('cavity', 48, 12, 4.0, 2500.0, 0.6, cg_tol=1e-13),
               ensemble.member('channel', 20, 16, 1.25, 120.0, 1.4, cg_tol=1e-13)]
    nsteps = 5
    B = host.Batch(members)
    B.input_check(1)
    B.step(nsteps)
    for k, m in enumerate(members):
        Vb, pb = B.get_state(k)
        Vo, po, _, _ = O.run(m, nsteps=nsteps)
        err = np.abs(Vb - Vo).max() / np.abs(Vo).max()
        if m.parameters()['kind'] == 'cavity':
            pb = pb - pb.mean(); po = po - po.mean()
        perr = np.abs(pb - po).max() / np.abs(po).max()
        _say(k, m.parameters()['kind'], f'V {err:.2e} p {perr:.2e}')
        assert err < 1e-11 and perr < 1e-10
    tot, mx, cells = B.stats()
    assert tot > 0 and cells == sum(s[4] * s[5] for s in B.sizes)
    B.close()


def test_ensemble_generator_sizes():
    ms = ensemble.members(64, seed=1)
    sizes = [(m.parameters()['Nx'], m.parameters()['Ny']) for m in ms]
    assert all(48 <= max(s) <= 512 and min(s) >= 8 for s in sizes)
    assert {m.parameters()['kind'] for m in ms} == {'cavity', 'channel'}

"""The oracle's time-dependent boundary values (BC_unsteady = 1), pinned on
the reference's TG_temporal case (case_files/TG_temporal/*.m).

- The reference-held solution results/TG_temporal/TG_Re100_N20_dt1e-3_Dir.mat
  (V, p at t = 1 on the 20x20 mesh, dt = 1e-3, copied under tests/golden/) is
  reproduced to round-off by RK44.  It pins the stage boundary values of
  time_ERK.m:62-64,106-109 and F.m:27-29 through V, and ydM of
  pressure_additional_solve.m:8-21 / set_bc_vectors.m:120-146 through p.
- Planted mistakes (ydM left out, boundary values frozen at t = 0) move the
  run far from that solution, so the comparison is sharp for both.
- RK44 converges with order 4 in time and the scheme with order 2 in space
  against the exact decaying vortex that the boundary functions sample.
"""
import os

import numpy as np
import pytest
from scipy.io import loadmat

from conftest import GOLDEN
from ins2d_amd.cases import load
from oracle import ins2d_oracle as O

FIXTURE = os.path.join(GOLDEN, 'TG_temporal', 'TG_Re100_N20_dt1e-3_Dir.mat')


def fixture():
    m = loadmat(FIXTURE)
    return m['V'].ravel(), m['p'].ravel()


def run(case):
    V, p, rows, opts = O.run(case)
    return V, p, rows[-1][2], opts


def test_fixture_reproduced_to_round_off():
    Vf, pf = fixture()
    V, p, t, _ = run(load('TG_temporal', dt=1e-3))
    assert abs(t - 1.0) < 1e-12
    assert np.max(np.abs(V - Vf)) < 1e-12
    assert np.max(np.abs(p - pf)) < 1e-12


def test_fixture_at_a_coarser_step():
    """dt = 0.01: the RK44 time error is far below the tolerance."""
    Vf, pf = fixture()
    V, p, _, _ = run(load('TG_temporal', dt=0.01))
    assert np.max(np.abs(V - Vf)) < 1e-9
    assert np.max(np.abs(p - pf)) < 1e-9


def test_planted_mistakes_are_seen():
    Vf, pf = fixture()
    # ydM left out of the additional pressure solve
    case = load('TG_temporal', dt=0.01)
    case.dudtBC = lambda x, y, t, o: np.zeros(len(np.atleast_1d(x)) * len(np.atleast_1d(y)))
    case.dvdtBC = case.dudtBC
    V, p, _, _ = run(case)
    assert np.max(np.abs(p - pf)) > 1e-4
    # boundary values frozen at t = 0 (treated as steady)
    case = load('TG_temporal', dt=0.01)
    u0, v0 = case.uBC, case.vBC
    case.uBC = lambda x, y, t, o: u0(x, y, 0.0, o)
    case.vBC = lambda x, y, t, o: v0(x, y, 0.0, o)
    V, p, _, _ = run(case)
    assert np.max(np.abs(V - Vf)) > 1e-3


def test_rk44_time_order_four():
    Vr, pr, _, _ = run(load('TG_temporal', dt=1 / 320))
    err = []
    for dt in (0.1, 0.05, 0.025):
        V, p, _, _ = run(load('TG_temporal', dt=dt))
        err.append((np.max(np.abs(V - Vr)), np.max(np.abs((p - p.mean()) - (pr - pr.mean())))))
    for a, b in zip(err, err[1:]):
        assert a[0] / b[0] > 14 and a[1] / b[1] > 14   # 2^4 = 16


@pytest.mark.parametrize('N', [20])
def test_space_order_two_against_the_exact_vortex(N):
    errs = []
    for n in (N, 2 * N):
        V, p, t, opts = run(load('TG_temporal', Nx=n, Ny=n, dt=0.005, t_end=0.5))
        u, v, pe = load('TG_temporal').module.exact(opts, t)
        e = V - np.r_[u.ravel(order='F'), v.ravel(order='F')]
        pe = pe.ravel(order='F')
        errs.append((np.sqrt(np.mean(e ** 2)), np.max(np.abs((p - p.mean()) - (pe - pe.mean())))))
    assert errs[0][0] / errs[1][0] > 3.0   # velocity, 2^2 = 4 asymptotically
    assert errs[0][1] / errs[1][1] > 3.5   # pressure

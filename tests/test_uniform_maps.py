"""Host side of the generic kernels' closed form (ins2d_amd/csrc/uniform.hpp),
on the CPU: tests/native/uniform_check.cpp builds operator plans and runs the
classification that decides which points take the closed form of the
periodic kernels instead of the kron-factored operators.

On a uniform all-wall cavity exactly the interior points (one point away
from every wall in the direction of each component's boundary closure) are
marked, and the grid-line offsets are those of the reference layout
(operator_mesh.m:185-320: u on x(2:end-1), v on y(2:end-1)); a channel with
a 'pres' outflow and a periodic direction keep their interiors; stretched
grids (the closed form with per-index spacings) mark their interiors too.  The device kernels are checked against the oracle by the
GPU parity tests; this pins the classification itself."""
import os
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope='module')
def rows(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp('uni') / 'uniform_check')
    subprocess.run(['/opt/rocm/bin/hipcc', '-O1', '-std=c++17', '-Xclang', '-target-feature', '-Xclang',
                    '-packed-fp32-ops', '-I', os.path.join(ROOT, 'include'), '-o', exe,
                    os.path.join(ROOT, 'tests/native/uniform_check.cpp'), os.path.join(ROOT, 'ins2d_amd/csrc/ops1d.cpp')],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.strip().splitlines():
        name, *v = line.split()
        res[name] = [int(x) for x in v]
    return res


def test_cavity_marks_exactly_the_interior(rows):
    ou, ov, nu, nv, su, sv, nux, nvy = rows['cavity']
    assert (ou, ov) == (1, 1)
    # u: 63 x 48 points, interior 61 x 46; v: 64 x 47, interior 62 x 45
    assert (nu, nv) == (63 * 48, 64 * 47)
    assert su == 61 * 46 and sv == 62 * 45


def test_channel_and_periodic_keep_their_interiors(rows):
    for name in ('channel', 'periodic_x'):
        ou, ov, nu, nv, su, sv, nux, nvy = rows[name]
        assert 0.85 * nu < su < nu and 0.85 * nv < sv < nv


def test_stretched_grids_mark_their_interior(rows):
    """The closed form with per-index spacings (stretched tensor grids, as the
    backward-facing step's tanh mesh) reproduces the generic operators at the
    interior points: a cavity stretched in y marks exactly its interior, a
    channel stretched in x and y its interior minus the outflow column."""
    ou, ov, nu, nv, su, sv, nux, nvy = rows['stretched']
    assert (nux, nvy) == (47, 39)
    assert su == 45 * 38 and sv == 46 * 37
    ou, ov, nu, nv, su, sv, nux, nvy = rows['stretched_xy']
    assert 0.8 * nu < su < nu and 0.8 * nv < sv < nv

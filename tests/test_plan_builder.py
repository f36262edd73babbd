"""The GPU library's host-side operator builder (ins2d_amd/csrc/ops1d.cpp)
against the CPU oracle's assembled operators, on the CPU.

tests/native/plan_dump.cpp links ops1d.cpp and prints every 1-D factor and
boundary term; here they are turned back into 2-D operators with kron() and
compared with the matrices the oracle assembles the reference's way, for
every boundary-kind combination the path supports, on uniform and stretched
grids, with non-zero boundary data.  The device kernels apply exactly these
factors, so this pins the operator data the GPU uses without a GPU."""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401

from conftest import ROOT
from ins2d_amd.cases import load
from ins2d_amd.host import BC_KIND, bc_samples
from oracle import ins2d_oracle as O


@pytest.fixture(scope='module')
def dumper(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp('plan') / 'plan_dump')
    subprocess.run(['g++', '-O1', '-std=c++17', '-o', exe, os.path.join(ROOT, 'tests/native/plan_dump.cpp'),
                    os.path.join(ROOT, 'ins2d_amd/csrc/ops1d.cpp'), os.path.join(ROOT, 'ins2d_amd/csrc/fdm_host.cpp')],
                   check=True)
    return exe


def _ell(o):
    K = o['K']
    rows = np.repeat(np.arange(o['nrows']), K)
    return sp.csr_matrix((np.array(o['val']), (rows, np.array(o['col']))), shape=(o['nrows'], o['ncols']))


def _kron(k):
    return sp.kron(_ell(k['Y']), _ell(k['X']), format='csr')


def _yv(terms):
    if not terms:
        return 0.0
    return sum(np.kron(np.array(ay), np.array(ax)) for ax, ay in terms)


def _case(kind, Nx, Ny, stretch=1.0):
    base = load('LDC_unsteady', Nx=Nx, Ny=Ny, Re=37.0, sx=stretch, sy=stretch)
    bcs = {
        'ldc': dict(u=dict(left='dir', right='dir', low='dir', up='dir'), v=dict(left='dir', right='dir', low='dir', up='dir')),
        'channel': dict(u=dict(left='per', right='per', low='dir', up='dir'), v=dict(left='per', right='per', low='dir', up='dir')),
        'vchannel': dict(u=dict(left='dir', right='dir', low='per', up='per'), v=dict(left='dir', right='dir', low='per', up='per')),
        'outflow': dict(u=dict(left='dir', right='pres', low='dir', up='dir'), v=dict(left='dir', right='sym', low='dir', up='dir')),
        'inflow_pres': dict(u=dict(left='pres', right='dir', low='sym', up='dir'), v=dict(left='sym', right='dir', low='dir', up='pres')),
        'pres_pres': dict(u=dict(left='pres', right='pres', low='dir', up='sym'), v=dict(left='sym', right='sym', low='pres', up='pres')),
        'periodic': dict(u=dict(left='per', right='per', low='per', up='per'), v=dict(left='per', right='per', low='per', up='per')),
    }[kind]
    base.BCtype = lambda: dict(BC_unsteady=0, **bcs)
    # smooth non-zero boundary data everywhere (each BC only reads its sides)
    base.uBC = lambda x, y, t, o: np.ravel(np.add.outer(np.atleast_1d(x), 0.5 * np.atleast_1d(y)) ** 2 + 0.2 +
                                           0.1 * np.sin(7 * np.add.outer(np.atleast_1d(x), np.atleast_1d(y))), order='F')
    base.vBC = lambda x, y, t, o: np.ravel(np.cos(2 * np.add.outer(np.atleast_1d(x), -np.atleast_1d(y))) - 0.4, order='F')
    return base


def _dump(dumper, opts, case, fdm=False):
    g = opts.grid
    bct = case.BCtype()
    lines = [f'{g.Nx} {g.Ny}', ' '.join('%.17g' % v for v in g.x), ' '.join('%.17g' % v for v in g.y),
             ' '.join(str(BC_KIND[bct[c][s]]) for c in 'uv' for s in ('left', 'right', 'low', 'up')),
             '%.17g' % opts.fluid.Re]
    for s in bc_samples(case, opts, 0.0):
        lines.append(str(len(s)) + ' ' + ' '.join('%.17g' % v for v in s))
    out = subprocess.run([dumper] + (['--fdm'] if fdm else []), input='\n'.join(lines), capture_output=True,
                         text=True, check=True).stdout
    d = json.loads(out)
    assert 'error' not in d, d
    return d


def _plan_F(d, V, p, comp):
    C = d[comp]
    Nu = d['u']['nx'] * d['u']['ny']
    u, v = V[:Nu], V[Nu:]
    own = u if comp == 'u' else v
    acc = 0.0
    for t in C['conv']:
        c = u if t['icomp'] == 0 else v
        face = (_kron(t['I']) @ c + _yv(t['yI'])) * (_kron(t['A']) @ own + _yv(t['yA']))
        acc = acc - t['scale'] * (_kron(t['D']) @ face)
    for t in C['diff']:
        acc = acc + _kron(t['D']) @ (t['sscale'] * (_kron(t['S']) @ own) + _yv(t['yS']))
    return acc - _kron(C['G']) @ p


CASES = [('ldc', 9, 7, 1.0), ('ldc', 8, 8, 1.07), ('channel', 10, 6, 1.0), ('vchannel', 7, 9, 1.0),
         ('outflow', 11, 6, 1.0), ('outflow', 10, 8, 1.05), ('inflow_pres', 8, 9, 1.0),
         ('pres_pres', 9, 8, 1.0), ('periodic', 8, 6, 1.0)]


@pytest.mark.parametrize('kind,Nx,Ny,stretch', CASES)
def test_plan_matches_oracle_operators(dumper, kind, Nx, Ny, stretch):
    case = _case(kind, Nx, Ny, stretch)
    opts = O.build(case)
    g, dsc = opts.grid, opts.discretization
    d = _dump(dumper, opts, case)
    assert d['sizes'] == [g.Nux_in, g.Nuy_in, g.Nvx_in, g.Nvy_in, g.Npx, g.Npy]
    np.testing.assert_array_equal(d['xin'], g.xin)
    np.testing.assert_array_equal(d['yin'], g.yin)
    rng = np.random.default_rng(11)
    V = rng.standard_normal(g.NV)
    p = rng.standard_normal(g.Np)
    _, Fo = O.F(V, V, p, 0.0, opts)
    Fp = np.r_[_plan_F(d, V, p, 'u'), _plan_F(d, V, p, 'v')]
    scale = np.abs(Fo).max()
    assert np.abs(Fp - Fo).max() <= 1e-13 * scale
    # divergence with boundary vector, gradient, volumes
    Mo = dsc.M @ V + dsc.yM
    Mp = _kron(d['u']['M']) @ V[:g.Nu] + _kron(d['v']['M']) @ V[g.Nu:] + _yv(d['u']['yM']) + _yv(d['v']['yM'])
    assert np.abs(Mp - Mo).max() <= 1e-13 * max(1.0, np.abs(Mo).max())
    Gp = np.r_[_kron(d['u']['G']) @ p, _kron(d['v']['G']) @ p]
    np.testing.assert_allclose(Gp, dsc.G @ p, atol=1e-14 * np.abs(dsc.G @ p).max())
    np.testing.assert_allclose(np.kron(d['u']['oy'], d['u']['ox']), g.Omu, rtol=1e-15)
    np.testing.assert_allclose(np.kron(d['v']['oy'], d['v']['ox']), g.Omv, rtol=1e-15)
    # pressure Laplacian and its diagonal
    A = _kron(d['L1']) + _kron(d['L2'])
    assert abs(A - dsc.A).max() <= 1e-13 * abs(dsc.A).max()
    np.testing.assert_allclose(d['Adiag'], dsc.A.diagonal(), rtol=1e-13)


def test_plan_rejects_bad_combination(dumper):
    case = _case('ldc', 6, 6)
    opts = O.build(case)
    g = opts.grid
    lines = [f'{g.Nx} {g.Ny}', ' '.join('%.17g' % v for v in g.x), ' '.join('%.17g' % v for v in g.y),
             '0 1 0 0 0 0 0 0', '1.0'] + ['0'] * 12
    out = subprocess.run([dumper] + [], input='\n'.join(lines), capture_output=True,
                         text=True, check=True).stdout
    assert 'error' in json.loads(out)


@pytest.mark.parametrize('kind,Nx,Ny,stretch', CASES)
def test_fast_diagonalization_factors(dumper, kind, Nx, Ny, stretch):
    """The direct solve P = Vx [-(Vx^T F Vy) / (lx_i + ly_j)] Vy^T built from
    the host eigen-decompositions (fdm_host.cpp) against a sparse LU solve of
    the oracle's A (modulo the constant where A is singular), plus the
    generalized eigen-equations and the diag(h)-orthonormality."""
    case = _case(kind, Nx, Ny, stretch)
    opts = O.build(case)
    g, A = opts.grid, opts.discretization.A
    d = _dump(dumper, opts, case, fdm=True)
    F = d['fdm']
    nx, ny = g.Npx, g.Npy
    Vx = np.array(F['Vx']).reshape(nx, nx, order='F')
    Vy = np.array(F['Vy']).reshape(ny, ny, order='F')
    lx, ly = np.array(F['lx']), np.array(F['ly'])
    Lx = _ell(d['L1']['X']).toarray()    # A = kron(diag(-hy), Lx) + kron(Ly, diag(-hx))
    Ly = _ell(d['L2']['Y']).toarray()
    hx, hy = np.asarray(g.hx), np.asarray(g.hy)
    assert np.abs(Lx @ Vx - (hx[:, None] * Vx) * lx).max() <= 1e-12 * np.abs(Lx).max()
    assert np.abs(Ly @ Vy - (hy[:, None] * Vy) * ly).max() <= 1e-12 * np.abs(Ly).max()
    np.testing.assert_allclose(Vx.T @ (hx[:, None] * Vx), np.eye(nx), atol=1e-12)
    np.testing.assert_allclose(Vy.T @ (hy[:, None] * Vy), np.eye(ny), atol=1e-12)
    singular = kind in ('ldc', 'channel', 'vchannel', 'periodic')
    assert (F['nx0'] >= 0) == singular
    rng = np.random.default_rng(5)
    f = rng.standard_normal(g.Np)
    if singular:
        f -= f.mean()
    Fm = f.reshape(nx, ny, order='F')
    G = Vx.T @ Fm @ Vy
    den = lx[:, None] + ly[None, :]
    if singular:
        den[F['nx0'], F['ny0']] = np.inf
    p = (Vx @ (-G / den) @ Vy.T).ravel(order='F')
    if singular:
        K = sp.bmat([[A, sp.csr_matrix(np.ones((g.Np, 1)))], [sp.csr_matrix(np.ones((1, g.Np))), None]], format='csc')
        pref = sp.linalg.spsolve(K, np.r_[f, 0.0])[:g.Np]
        p, pref = p - p.mean(), pref - pref.mean()
    else:
        pref = sp.linalg.spsolve(A.tocsc(), f)
    assert np.abs(p - pref).max() <= 1e-11 * np.abs(pref).max()
    assert np.abs(A @ p - f).max() <= 1e-11 * np.abs(f).max()


@pytest.mark.parametrize('kind,Nx,Ny,stretch', CASES)
def test_ic0_factor_matches_make_cholinc(dumper, kind, Nx, Ny, stretch):
    """The host IC(0) diagonal (fdm_host.cpp build_ic0_host) against the
    oracle's restatement of make_cholinc.m on the assembled A."""
    case = _case(kind, Nx, Ny, stretch)
    opts = O.build(case)
    d = _dump(dumper, opts, case, fdm=True)
    assert 'ic0' in d, d.get('ic0_error')
    Dref = O.make_cholinc_diag(opts.discretization.A)
    np.testing.assert_allclose(d['ic0']['D'], Dref, rtol=1e-12, atol=0)

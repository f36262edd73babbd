"""CPU oracle for the INS2D unsteady explicit-RK hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (``ins2d_amd``) never
imports it and fails loudly when its HIP library is missing.

This is a restatement, in numpy/scipy, of the reference MATLAB algorithm: it
assembles the *same* sparse kron/boundary-closure operators the reference
builds and applies them the same way.  Every function cites the reference file
and lines it follows (paths relative to the reference root).

Parity pins (what checks this module against something other than itself):
  * reference-held whole-run fixtures: results/LDC_unsteady_1.000e+02_20x20_0
    and _100x100_0 (convergence.txt rows and 17-digit centreline profiles),
    results/shear_layer01/02/03 and shear_layer_snapshotdata convergence logs
    -> tests/test_oracle_fixtures.py;
  * closed forms / invariants: G = -M^T, zero row sums of A, skew-symmetry of
    the convection operator, the 2nd-order FFT symbol equal to the eigenvalues
    of the assembled A, exact discrete divergence of V = W^T psi, IC(0) PCG vs a
    direct solve -> tests/test_oracle_invariants.py.
  * 4th-order operators with wall boundaries (BC_*3 closures with 'dir'): only
    the periodic branch is pinned (by symbols and invariants); the wall branch
    is restated but parity unpinned.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# threads of the FFT pressure solve (scipy.fft workers); the benchmark's CPU
# leg sets it to the host's core count, everything else runs it single-threaded
FFT_WORKERS = 1

# ----------------------------------------------------------------------------
# MATLAB sparse helpers
# ----------------------------------------------------------------------------


def mspdiags(cols, offsets, m, n):
    """MATLAB ``spdiags(B, d, m, n)`` with MATLAB's indexing convention.

    For m >= n the diagonal values are taken by column index, for m < n by row
    index (MATLAB documentation of spdiags).  ``cols`` is a list of 1-D arrays
    (the columns of B) or scalars.
    """
    rows, cs, vals = [], [], []
    for B, d in zip(cols, offsets):
        if m >= n:
            j = np.arange(n)
            i = j - d
        else:
            i = np.arange(m)
            j = i + d
        ok = (i >= 0) & (i < m) & (j >= 0) & (j < n)
        idx = j if m >= n else i
        Bv = np.broadcast_to(np.asarray(B, dtype=float), (max(m, n) + 8,)) if np.ndim(B) == 0 else np.asarray(B, float)
        ok &= idx < len(Bv)
        rows.append(i[ok]); cs.append(j[ok]); vals.append(Bv[idx[ok]])
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cs))), shape=(m, n))
    return A.tocsr()


def speye(n):
    return sp.identity(n, format="csr")


def sdiag(v):
    v = np.asarray(v, float)
    return sp.diags(v, 0, shape=(len(v), len(v)), format="csr")


def kron(a, b):
    return sp.kron(a, b, format="csr")


# ----------------------------------------------------------------------------
# 1-D boundary closures (spatial/boundaryconditions/proposed/*.m)
# ----------------------------------------------------------------------------


def _closure(Nt, Nin, Nb, bin_offset, bb_entries, bbc_entries, y1, y2):
    """u_ext = B1D u_in + Btemp ybc with Btemp = Bb (Bbc Bb)^-1 and
    B1D = Bin - Btemp Bbc Bin (BC_general.m:1-7,124-125)."""
    Bin = mspdiags([np.ones(Nt)], [bin_offset], Nt, Nin)
    Bb = sp.lil_matrix((Nt, Nb))
    for (i, j, v) in bb_entries:
        Bb[i, j] = v
    Bbc = sp.lil_matrix((Nb, Nt))
    for (i, j, v) in bbc_entries:
        Bbc[i, j] = v
    Bb = Bb.tocsr(); Bbc = Bbc.tocsr()
    S = (Bbc @ Bb).toarray()
    Btemp = sp.csr_matrix(Bb @ np.linalg.inv(S))
    B1D = (Bin - Btemp @ (Bbc @ Bin)).tocsr()
    B1D.eliminate_zeros()
    return SimpleNamespace(B1D=B1D, Btemp=Btemp, ybc1=np.asarray(y1, float), ybc2=np.asarray(y2, float))


def _nb0(Nt):
    return SimpleNamespace(B1D=speye(Nt), Btemp=sp.csr_matrix((Nt, 2)), ybc1=np.zeros(2), ybc2=np.zeros(2))


def BC_general(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """spatial/boundaryconditions/proposed/BC_general.m:1-130 (ghosts on grid lines)."""
    assert Nt == Nin + Nb
    if Nb == 0:
        return _nb0(Nt)
    y1 = np.zeros(Nb); y2 = np.zeros(Nb)
    bbc = {}
    if Nb == 2:
        bb = [(0, 0, 1.0), (Nt - 1, 1, 1.0)]
        off = -1
        if BC1 == 'dir':
            bbc[(0, 0)] = 1.0; y1[0] = 1.0
        elif BC1 == 'pres':
            bbc[(0, 0)] = -1.0; bbc[(0, 1)] = 1.0; y1[0] = 2 * h1
        elif BC1 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 2)] = 1.0; bbc[(1, 1)] = -1.0; bbc[(1, Nt - 1)] = 1.0
        else:
            raise ValueError('not implemented')
        if BC2 == 'dir':
            bbc[(1, Nt - 1)] = 1.0; y2[1] = 1.0
        elif BC2 == 'pres':
            bbc[(1, Nt - 2)] = -1.0; bbc[(1, Nt - 1)] = 1.0; y2[1] = 2 * h2
        elif BC2 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 2)] = 1.0; bbc[(1, 1)] = -1.0; bbc[(1, Nt - 1)] = 1.0
        else:
            raise ValueError('not implemented')
    elif Nb == 1:
        bb = []
        off = -1
        if BC1 == 'dir':
            bbc[(0, 0)] = 1.0; y1[0] = 1.0; bb.append((0, 0, 1.0))
        elif BC1 == 'pres':
            off = 0
        elif BC1 == 'per':
            off = 0; bbc[(0, 0)] = -1.0; bbc[(0, Nt - 1)] = 1.0; bb.append((Nt - 1, 0, 1.0))
        else:
            raise ValueError('not implemented')
        if BC2 == 'dir':
            bbc[(0, Nt - 1)] = 1.0; y2[0] = 1.0; bb.append((Nt - 1, 0, 1.0))
        elif BC2 == 'pres':
            pass
        elif BC2 == 'per':
            off = 0; bbc[(0, 0)] = -1.0; bbc[(0, Nt - 1)] = 1.0; bb.append((Nt - 1, 0, 1.0))
        else:
            raise ValueError('not implemented')
    else:
        raise ValueError('unsupported Nb')
    return _closure(Nt, Nin, Nb, off, bb, [(i, j, v) for (i, j), v in bbc.items()], y1, y2)


def BC_general_stag(Nt, Nin, Nb, BC1, BC2, h1, h2, diff_variant=False):
    """spatial/boundaryconditions/proposed/BC_general_stag.m:1-138 (ghosts on
    staggered lines; Dirichlet by mirror ½(u0+u1)=ub).  ``diff_variant`` selects
    BC_diff_stag.m:1-160, whose Nb=2 branch has identical closures."""
    assert Nt == Nin + Nb
    if Nb == 0:
        return _nb0(Nt)
    y1 = np.zeros(Nb); y2 = np.zeros(Nb)
    bbc = {}
    if Nb == 2:
        bb = [(0, 0, 1.0), (Nt - 1, 1, 1.0)]
        off = -1
        if BC1 == 'dir':
            bbc[(0, 0)] = 0.5; bbc[(0, 1)] = 0.5; y1[0] = 1.0
        elif BC1 == 'sym':
            bbc[(0, 0)] = -1.0; bbc[(0, 1)] = 1.0; y1[0] = h1
        elif BC1 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 2)] = 1.0; bbc[(1, 1)] = -1.0; bbc[(1, Nt - 1)] = 1.0
        else:
            raise ValueError('not implemented')
        if BC2 == 'dir':
            bbc[(1, Nt - 2)] = 0.5; bbc[(1, Nt - 1)] = 0.5; y2[1] = 1.0
        elif BC2 == 'sym':
            bbc[(1, Nt - 2)] = -1.0; bbc[(1, Nt - 1)] = 1.0; y2[1] = h2
        elif BC2 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 2)] = 1.0; bbc[(1, 1)] = -1.0; bbc[(1, Nt - 1)] = 1.0
        else:
            raise ValueError('not implemented')
    elif Nb == 1 and not diff_variant:
        bb = []
        off = -1
        if BC1 == 'dir':
            bbc[(0, 0)] = 0.5; bbc[(0, 1)] = 0.5; y1[0] = 1.0; bb.append((0, 0, 1.0))
        elif BC1 == 'sym':
            bbc[(0, 0)] = -1.0; bbc[(0, 1)] = 1.0; y1[0] = h1
        elif BC1 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 1)] = 1.0; bb.append((0, 0, 1.0))
        else:
            raise ValueError('not implemented')
        if BC2 == 'dir':
            bbc[(0, Nt - 2)] = 0.5; bbc[(0, Nt - 1)] = 0.5; y2[0] = 1.0; bb.append((Nt - 1, 0, 1.0))
        elif BC2 == 'sym':
            bbc[(0, Nt - 2)] = -1.0; bbc[(0, Nt - 1)] = 1.0; y2[0] = h2
        elif BC2 == 'per':
            bbc[(0, 0)] = -1.0; bbc[(0, Nt - 1)] = 1.0; bb.append((0, 0, 1.0))
        else:
            raise ValueError('not implemented')
    else:
        raise ValueError('unsupported Nb')
    return _closure(Nt, Nin, Nb, off, bb, [(i, j, v) for (i, j), v in bbc.items()], y1, y2)


def BC_diff_stag(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """spatial/boundaryconditions/proposed/BC_diff_stag.m:1-160 ('zeroth order'
    mirror closure; same matrices as BC_general_stag for Nb=2)."""
    return BC_general_stag(Nt, Nin, Nb, BC1, BC2, h1, h2, diff_variant=True)


def _per_block(Nt, k):
    """Periodic closure rows for k ghost points at each end (BC_*3.m 'per')."""
    e = {}
    for r in range(k):
        e[(r, r)] = -1.0; e[(r, Nt - 2 * k + r)] = 1.0
        e[(k + r, k + r)] = -1.0; e[(k + r, Nt - k + r)] = 1.0
    return e


def _bc_wide(Nt, Nin, Nb, BC1, BC2, k, dir_lo, dir_hi, name):
    """Generic body of the 4th-order closures with k ghosts per side.

    dir_lo(e, y1) / dir_hi(e, y2) fill the Dirichlet rows."""
    assert Nt == Nin + Nb
    if Nb == 0:
        return _nb0(Nt)
    if Nb != 2 * k:
        raise ValueError(f'{name}: unsupported Nb={Nb}')
    y1 = np.zeros(Nb); y2 = np.zeros(Nb)
    bb = [(r, r, 1.0) for r in range(k)] + [(Nt - k + r, Nb - k + r, 1.0) for r in range(k)]
    e = {}
    if BC1 == 'per' or BC2 == 'per':
        if not (BC1 == 'per' and BC2 == 'per'):
            raise ValueError(f'{name}: mixed periodic')
        e.update(_per_block(Nt, k))
    else:
        if BC1 != 'dir' or BC2 != 'dir':
            raise ValueError(f'{name}: not implemented for {BC1}/{BC2}')
        dir_lo(e, y1); dir_hi(e, y2)
    return _closure(Nt, Nin, Nb, -k, bb, [(i, j, v) for (i, j), v in e.items()], y1, y2)


def BC_int2(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_int2.m (k=2): dir rows ½(u1+u3)=ub, u2=ub."""
    def lo(e, y):
        e[(0, 0)] = .5; e[(0, 2)] = .5; e[(1, 1)] = 1.; y[0] = 1; y[1] = 1
    def hi(e, y):
        n = Nt; b = Nb
        e[(b - 1, n - 1)] = .5; e[(b - 1, n - 3)] = .5; e[(b - 2, n - 2)] = 1.; y[b - 2] = 1; y[b - 1] = 1
    return _bc_wide(Nt, Nin, Nb, BC1, BC2, 2, lo, hi, 'BC_int2')


def BC_div2(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_div2.m (k=2): same closure rows as BC_int2."""
    return BC_int2(Nt, Nin, Nb, BC1, BC2, h1, h2)


def _dir3_sym(e, y, Nt, Nb, lo):
    # ½(u1+u5)=ub, ½(u2+u4)=ub, u3=ub   (BC_av3.m / BC_int3.m / BC_diff3.m 'dir')
    if lo:
        e[(0, 0)] = .5; e[(0, 4)] = .5; e[(1, 1)] = .5; e[(1, 3)] = .5; e[(2, 2)] = 1.; y[0:3] = 1
    else:
        n = Nt; b = Nb
        e[(b - 1, n - 1)] = .5; e[(b - 1, n - 5)] = .5; e[(b - 2, n - 2)] = .5; e[(b - 2, n - 4)] = .5
        e[(b - 3, n - 3)] = 1.; y[b - 3:b] = 1


def BC_av3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_av3.m (k=3)."""
    return _bc_wide(Nt, Nin, Nb, BC1, BC2, 3, lambda e, y: _dir3_sym(e, y, Nt, Nb, True),
                    lambda e, y: _dir3_sym(e, y, Nt, Nb, False), 'BC_av3')


def BC_int3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_int3.m (k=3): same rows as BC_av3."""
    return BC_av3(Nt, Nin, Nb, BC1, BC2, h1, h2)


def BC_diff3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_diff3.m (k=3): same rows as BC_av3."""
    return BC_av3(Nt, Nin, Nb, BC1, BC2, h1, h2)


def BC_av_stag3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_av_stag3.m (k=3): -u1+u6=0, ½(u2+u5)=ub, -u3+u4=0."""
    def lo(e, y):
        e[(0, 0)] = -1.; e[(0, 5)] = 1.; e[(1, 1)] = .5; e[(1, 4)] = .5; e[(2, 2)] = -1.; e[(2, 3)] = 1.; y[1] = 1
    def hi(e, y):
        n = Nt; b = Nb
        e[(b - 1, n - 1)] = -1.; e[(b - 1, n - 6)] = 1.; e[(b - 2, n - 2)] = .5; e[(b - 2, n - 5)] = .5
        e[(b - 3, n - 3)] = -1.; e[(b - 3, n - 4)] = 1.; y[b - 2] = 1
    return _bc_wide(Nt, Nin, Nb, BC1, BC2, 3, lo, hi, 'BC_av_stag3')


def BC_diff_stag3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_diff_stag3.m (k=3): mirror rows ½(u1+u6)=½(u2+u5)=½(u3+u4)=ub."""
    def lo(e, y):
        e[(0, 0)] = .5; e[(0, 5)] = .5; e[(1, 1)] = .5; e[(1, 4)] = .5; e[(2, 2)] = .5; e[(2, 3)] = .5; y[0:3] = 1
    def hi(e, y):
        n = Nt; b = Nb
        e[(b - 1, n - 1)] = .5; e[(b - 1, n - 6)] = .5; e[(b - 2, n - 2)] = .5; e[(b - 2, n - 5)] = .5
        e[(b - 3, n - 3)] = .5; e[(b - 3, n - 4)] = .5; y[b - 3:b] = 1
    return _bc_wide(Nt, Nin, Nb, BC1, BC2, 3, lo, hi, 'BC_diff_stag3')


def _mixed_stag(Nt, Nin, Nb, BC1, BC2, name):
    """BC_int_mixed_stag2.m / BC_int_mixed_stag3.m (k=2): Neumann-type rows
    -u1+u4=0, -u2+u3=0 for 'dir' (skew-symmetry)."""
    def lo(e, y):
        e[(0, 0)] = -1.; e[(0, 3)] = 1.; e[(1, 1)] = -1.; e[(1, 2)] = 1.
    def hi(e, y):
        n = Nt; b = Nb
        e[(b - 1, n - 1)] = -1.; e[(b - 1, n - 4)] = 1.; e[(b - 2, n - 2)] = -1.; e[(b - 2, n - 3)] = 1.
    return _bc_wide(Nt, Nin, Nb, BC1, BC2, 2, lo, hi, name)


def BC_int_mixed_stag2(Nt, Nin, Nb, BC1, BC2, h1, h2):
    return _mixed_stag(Nt, Nin, Nb, BC1, BC2, 'BC_int_mixed_stag2')


def BC_int_mixed_stag3(Nt, Nin, Nb, BC1, BC2, h1, h2):
    return _mixed_stag(Nt, Nin, Nb, BC1, BC2, 'BC_int_mixed_stag3')


def BC_int_mixed2(Nt, Nin, Nb, BC1, BC2, h1, h2):
    """BC_int_mixed2.m: dir/dir (k=2) or per/per (Nb=3, offset -1)."""
    assert Nt == Nin + Nb
    y1 = np.zeros(Nb); y2 = np.zeros(Nb)
    if BC1 == 'dir' and BC2 == 'dir':
        bb = [(0, 0, 1.), (1, 1, 1.), (Nt - 2, Nb - 2, 1.), (Nt - 1, Nb - 1, 1.)]
        e = {(0, 0): -1., (0, 2): 1., (1, 1): 1., (Nb - 1, Nt - 1): -1., (Nb - 1, Nt - 3): 1., (Nb - 2, Nt - 2): 1.}
        y1[1] = 1; y2[Nb - 2] = 1
        return _closure(Nt, Nin, Nb, -2, bb, [(i, j, v) for (i, j), v in e.items()], y1, y2)
    if BC1 == 'per' and BC2 == 'per':
        bb = [(0, 0, 1.), (Nt - 2, Nb - 2, 1.), (Nt - 1, Nb - 1, 1.)]
        e = {(0, 0): -1., (0, Nt - 3): 1., (Nb - 2, 1): -1., (Nb - 2, Nt - 2): 1., (Nb - 1, 2): -1., (Nb - 1, Nt - 1): 1.}
        return _closure(Nt, Nin, Nb, -1, bb, [(i, j, v) for (i, j), v in e.items()], y1, y2)
    raise ValueError('BC_int_mixed2: not implemented')


# ----------------------------------------------------------------------------
# mesh (preprocessing/mesh_generation.m, libs/nonuniform_grid.m)
# ----------------------------------------------------------------------------


def nonuniform_grid(deltaz, z_low, z_up, sz):
    """libs/nonuniform_grid.m:1-64 (uniform branch and geometric stretching)."""
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
    sz = s[-1]
    zz = np.zeros(n + 1)
    zz[0] = z_low
    for k in range(n):
        zz[k + 1] = zz[k] + deltaz * sz ** k
    return zz


def mesh_generation(opts, x, y):
    """preprocessing/mesh_generation.m:17-47."""
    g = opts.grid
    x = np.asarray(x, float); y = np.asarray(y, float)
    g.x, g.y = x, y
    g.xp = (x[:-1] + x[1:]) / 2
    g.yp = (y[:-1] + y[1:]) / 2
    g.hx = np.diff(x); g.hy = np.diff(y)
    g.Nx = len(g.hx); g.Ny = len(g.hy)
    for h, name in ((g.hx, 'gx'), (g.hy, 'gy')):
        N = len(h)
        gg = np.zeros(N + 1)
        gg[0] = h[0] / 2
        gg[1:N] = (h[:-1] + h[1:]) / 2
        gg[N] = h[-1] / 2
        setattr(g, name, gg)


# ----------------------------------------------------------------------------
# operators (spatial/operators/*.m)
# ----------------------------------------------------------------------------


def operator_mesh(opts):
    """spatial/operators/operator_mesh.m:1-531."""
    BC = opts.BC; g = opts.grid; d = opts.discretization
    order4 = d.order4; alfa = d.alfa
    Nx, Ny = g.Nx, g.Ny
    x, y, hx, hy, gx, gy, xp, yp = g.x, g.y, g.hx, g.hy, g.gx, g.gy, g.xp, g.yp
    g.Npx, g.Npy = Nx, Ny
    g.Np = Nx * Ny
    Nux_in = Nx + 1
    if BC.u.left in ('dir', 'sym'): Nux_in -= 1
    if BC.u.right in ('dir', 'sym'): Nux_in -= 1
    if BC.u.left == 'per' and BC.u.right == 'per': Nux_in -= 1
    Nuy_in = Ny
    Nvx_in = Nx
    Nvy_in = Ny + 1
    if BC.v.low in ('dir', 'sym'): Nvy_in -= 1
    if BC.v.up in ('dir', 'sym'): Nvy_in -= 1
    if BC.v.low == 'per' and BC.v.up == 'per': Nvy_in -= 1
    g.Nux_in, g.Nux_b, g.Nux_t = Nux_in, 2, Nux_in + 2
    g.Nuy_in, g.Nuy_b, g.Nuy_t = Nuy_in, 2, Nuy_in + 2
    g.Nvx_in, g.Nvx_b, g.Nvx_t = Nvx_in, 2, Nvx_in + 2
    g.Nvy_in, g.Nvy_b, g.Nvy_t = Nvy_in, 2, Nvy_in + 2
    g.Nu = Nux_in * Nuy_in
    g.Nv = Nvx_in * Nvy_in
    g.NV = g.Nu + g.Nv

    if order4:
        def three(h, per):
            N = len(h)
            h3 = np.zeros(N)
            h3[1:-1] = h[:-2] + h[1:-1] + h[2:]
            if per:
                h3[0] = h[-1] + h[0] + h[1]; h3[-1] = h[-2] + h[-1] + h[0]
            else:
                h3[0] = 2 * h[0] + h[1]; h3[-1] = h[-2] + 2 * h[-1]
            return h3

        def gthree(gg, N, per):
            g3 = np.zeros(N + 1)
            # gx3(3:Nx-1) = gx(2:end-3)+gx(3:end-2)+gx(4:end-1)  (1-based)
            g3[2:N - 1] = gg[1:-3] + gg[2:-2] + gg[3:-1]
            if per:
                g3[0] = gg[-2] + gg[-1] + gg[0] + gg[1]
                g3[1] = gg[-1] + gg[0] + gg[1] + gg[2]
                g3[-2] = gg[-3] + gg[-2] + gg[-1] + gg[0]
                g3[-1] = gg[-2] + gg[-1] + gg[0] + gg[1]
            else:
                g3[0] = 2 * gg[0] + 2 * gg[1]
                g3[1] = 2 * gg[0] + gg[1] + gg[2]
                g3[-2] = 2 * gg[-1] + gg[-2] + gg[-3]
                g3[-1] = 2 * gg[-1] + 2 * gg[-2]
            return g3

        perx = BC.u.left == 'per' and BC.u.right == 'per'
        pery = BC.v.low == 'per' and BC.v.up == 'per'
        hx3 = three(hx, perx); hy3 = three(hy, pery)
        gx3 = gthree(gx, Nx, perx); gy3 = gthree(gy, Ny, pery)
        g.hx3, g.hy3, g.hxi3, g.hyi3 = hx3, hy3, hx3.copy(), hy3.copy()

    gxd = gx.copy(); gxd[0] = hx[0]; gxd[-1] = hx[-1]
    hxi = hx.copy()
    L, R = BC.u.left, BC.u.right
    if L == 'dir' and R == 'dir':
        xin = x[1:-1]; hxd = hx.copy(); gxi = gx[1:-1]; diagpos = 1
        if order4:
            hxd3 = np.r_[hx3[0], hx3, hx3[-1]]; hxd13 = np.r_[hx[0], hx, hx[-1]]
            gxd3 = np.r_[2 * gx[0] + gx[1] + gx[2], gx3, 2 * gx[-1] + gx[-2] + gx[-3]]
            gxd13 = np.r_[gx[1], 2 * gx[0], gx[1:-1], 2 * gx[-1], gx[-2]]
            gxi3 = gx3[1:-1]
    elif L == 'dir' and R == 'pres':
        xin = x[1:]; hxd = np.r_[hx, hx[-1]]; gxi = gx[1:]; diagpos = 1
    elif L == 'pres' and R == 'dir':
        xin = x[:-1]; hxd = np.r_[hx[0], hx]; gxi = gx[:-1]; diagpos = 0
    elif L == 'pres' and R == 'pres':
        xin = x.copy(); hxd = np.r_[hx[0], hx, hx[-1]]; gxi = gx.copy(); diagpos = 0
    elif L == 'per' and R == 'per':
        xin = x[:-1]; hxd = np.r_[hx[-1], hx]; gxi = np.r_[gx[0] + gx[-1], gx[1:-1]]
        gxd[0] = (hx[0] + hx[-1]) / 2; gxd[-1] = (hx[0] + hx[-1]) / 2; diagpos = 0
        if order4:
            hxd3 = np.r_[hx3[-2], hx3[-1], hx3, hx3[0]]
            hxd13 = np.r_[hx[-2], hx[-1], hx, hx[0]]
            gxd3 = np.r_[gx3[-2], gx3, gx3[1]]
            gxd13 = np.r_[gx[-2], gx[0] + gx[-1], gx[1:-1], gx[-1] + gx[0], gx[1]]
            gxi3 = gx3[:-1]
    else:
        raise ValueError('unsupported u left/right BC combination')
    g.Bvux = mspdiags([np.ones(g.Nvx_t - 1)], [diagpos], Nux_in, g.Nvx_t - 1)

    gyd = gy.copy(); gyd[0] = hy[0]; gyd[-1] = hy[-1]
    hyi = hy.copy()
    Lo, Up = BC.v.low, BC.v.up
    if Lo == 'dir' and Up == 'dir':
        yin = y[1:-1]; hyd = hy.copy(); gyi = gy[1:-1]; diagpos = 1
        if order4:
            hyd3 = np.r_[hy3[0], hy3, hy3[-1]]; hyd13 = np.r_[hy[0], hy, hy[-1]]
            gyd3 = np.r_[2 * gy[0] + gy[1] + gy[2], gy3, 2 * gy[-1] + gy[-2] + gy[-3]]
            gyd13 = np.r_[gy[1], 2 * gy[0], gy[1:-1], 2 * gy[-1], gy[-2]]
            gyi3 = gy3[1:-1]
    elif Lo == 'dir' and Up == 'pres':
        yin = y[1:]; hyd = np.r_[hy, hy[-1]]; gyi = gy[1:]; diagpos = 1
    elif Lo == 'pres' and Up == 'dir':
        yin = y[:-1]; hyd = np.r_[hy[0], hy]; gyi = gy[:-1]; diagpos = 0
    elif Lo == 'pres' and Up == 'pres':
        yin = y.copy(); hyd = np.r_[hy[0], hy, hy[-1]]; gyi = gy.copy(); diagpos = 0
    elif Lo == 'per' and Up == 'per':
        yin = y[:-1]; hyd = np.r_[hy[-1], hy]; gyi = np.r_[gy[0] + gy[-1], gy[1:-1]]
        gyd[0] = (hy[0] + hy[-1]) / 2; gyd[-1] = (hy[0] + hy[-1]) / 2; diagpos = 0
        if order4:
            hyd3 = np.r_[hy3[-2], hy3[-1], hy3, hy3[0]]
            hyd13 = np.r_[hy[-2], hy[-1], hy, hy[0]]
            gyd3 = np.r_[gy3[-2], gy3, gy3[1]]
            gyd13 = np.r_[gy[-2], gy[0] + gy[-1], gy[1:-1], gy[-1] + gy[0], gy[1]]
            gyi3 = gy3[:-1]
    else:
        raise ValueError('unsupported v low/up BC combination')
    g.Buvy = mspdiags([np.ones(g.Nuy_t - 1)], [diagpos], Nvy_in, g.Nuy_t - 1)

    g.Omp = np.kron(hyi, hxi)
    g.Omu = np.kron(hyi, gxi)
    g.Omv = np.kron(gyi, hxi)
    g.Omux = np.kron(hyi, hxd)
    g.Omuy = np.kron(gyd, gxi)
    g.Omvx = np.kron(gyi, gxd)
    g.Omvy = np.kron(hyd, hxi)
    if order4:
        g.Omux1 = np.kron(hyi, hxd13); g.Omuy1 = np.kron(gyd13, gxi)
        g.Omvx1 = np.kron(gyi, gxd13); g.Omvy1 = np.kron(hyd13, hxi)
        g.Omux3 = np.kron(hy3, hxd3); g.Omuy3 = np.kron(gyd3, gxi3)
        g.Omvx3 = np.kron(gyi3, gxd3); g.Omvy3 = np.kron(hyd3, hx3)
        Omu3 = np.kron(hy3, gxi3); Omv3 = np.kron(gyi3, hx3)
        g.Omu = alfa * g.Omu - Omu3
        g.Omv = alfa * g.Omv - Omv3
        g.Omux = alfa * g.Omux1 - g.Omux3
        g.Omuy = alfa * g.Omuy1 - g.Omuy3
        g.Omvx = alfa * g.Omvx1 - g.Omvx3
        g.Omvy = alfa * g.Omvy1 - g.Omvy3
        g.gxi3, g.gyi3 = gxi3, gyi3
        g.hxd13, g.hxd3, g.hyd13, g.hyd3 = hxd13, hxd3, hyd13, hyd3
        g.gxd13, g.gxd3, g.gyd13, g.gyd3 = gxd13, gxd3, gyd13, gyd3
    g.Om = np.r_[g.Omu, g.Omv]
    g.Om_inv = 1.0 / g.Om
    g.Omu_inv = 1.0 / g.Omu; g.Omv_inv = 1.0 / g.Omv
    g.Omp_inv = 1.0 / g.Omp
    g.hxi, g.hyi, g.hxd, g.hyd = hxi, hyi, hxd, hyd
    g.gxi, g.gyi, g.gxd, g.gyd = gxi, gyi, gxd, gyd
    g.xin, g.yin = xin, yin
    # metrics for initialisation (operator_mesh.m:401-414), column-major 2-D arrays
    g.xu = np.outer(xin, np.ones(Nuy_in)); g.yu = np.outer(np.ones(Nux_in), yp)
    g.xv = np.outer(xp, np.ones(Nvy_in)); g.yv = np.outer(np.ones(Nvx_in), yin)
    g.xpp = np.outer(xp, np.ones(Ny)); g.ypp = np.outer(np.ones(Nx), yp)


def operator_averaging(opts):
    """spatial/operators/operator_averaging.m:1-158."""
    BC = opts.BC; g = opts.grid; d = opts.discretization
    hx, hy = g.hx, g.hy
    w = 0.5

    def A1(n):
        return mspdiags([w * np.ones(n), w * np.ones(n)], [0, 1], n - 1, n)

    b = BC_general(g.Nux_t, g.Nux_in, g.Nux_b, BC.u.left, BC.u.right, hx[0], hx[-1])
    d.Au_ux = kron(speye(g.Nuy_in), A1(g.Nux_t) @ b.B1D); b.Bbc = kron(speye(g.Nuy_in), A1(g.Nux_t) @ b.Btemp); d.Au_ux_BC = b
    b = BC_general_stag(g.Nuy_t, g.Nuy_in, g.Nuy_b, BC.u.low, BC.u.up, hy[0], hy[-1])
    d.Au_uy = kron(A1(g.Nuy_t) @ b.B1D, speye(g.Nux_in)); b.Bbc = kron(A1(g.Nuy_t) @ b.Btemp, speye(g.Nux_in)); d.Au_uy_BC = b
    b = BC_general_stag(g.Nvx_t, g.Nvx_in, g.Nvx_b, BC.v.left, BC.v.right, hx[0], hx[-1])
    d.Av_vx = kron(speye(g.Nvy_in), A1(g.Nvx_t) @ b.B1D); b.Bbc = kron(speye(g.Nvy_in), A1(g.Nvx_t) @ b.Btemp); d.Av_vx_BC = b
    b = BC_general(g.Nvy_t, g.Nvy_in, g.Nvy_b, BC.v.low, BC.v.up, hy[0], hy[-1])
    d.Av_vy = kron(A1(g.Nvy_t) @ b.B1D, speye(g.Nvx_in)); b.Bbc = kron(A1(g.Nvy_t) @ b.Btemp, speye(g.Nvx_in)); d.Av_vy_BC = b

    if d.order4:
        def A3(nin, nt):
            return mspdiags([w * np.ones(nt + 4), w * np.ones(nt + 4)], [0, 3], nin + 3, nt + 4)
        b = BC_av3(g.Nux_t + 4, g.Nux_in, g.Nux_t + 4 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
        A = A3(g.Nux_in, g.Nux_t)
        d.Au_ux3 = kron(speye(g.Nuy_in), A @ b.B1D); b.Bbc = kron(speye(g.Nuy_in), A @ b.Btemp); d.Au_ux_BC3 = b
        b = BC_av_stag3(g.Nuy_t + 4, g.Nuy_in, g.Nuy_t + 4 - g.Nuy_in, BC.u.low, BC.u.up, hy[0], hy[-1])
        A = A3(g.Nuy_in, g.Nuy_t)
        d.Au_uy3 = kron(A @ b.B1D, speye(g.Nux_in)); b.Bbc = kron(A @ b.Btemp, speye(g.Nux_in)); d.Au_uy_BC3 = b
        b = BC_av_stag3(g.Nvx_t + 4, g.Nvx_in, g.Nvx_t + 4 - g.Nvx_in, BC.v.left, BC.v.right, hx[0], hx[-1])
        A = A3(g.Nvx_in, g.Nvx_t)
        d.Av_vx3 = kron(speye(g.Nvy_in), A @ b.B1D); b.Bbc = kron(speye(g.Nvy_in), A @ b.Btemp); d.Av_vx_BC3 = b
        b = BC_av3(g.Nvy_t + 4, g.Nvy_in, g.Nvy_t + 4 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
        A = A3(g.Nvy_in, g.Nvy_t)
        d.Av_vy3 = kron(A @ b.B1D, speye(g.Nvx_in)); b.Bbc = kron(A @ b.Btemp, speye(g.Nvx_in)); d.Av_vy_BC3 = b


def operator_interpolation(opts):
    """spatial/operators/operator_interpolation.m:1-367."""
    BC = opts.BC; g = opts.grid; d = opts.discretization
    Nx, Ny = g.Nx, g.Ny
    hx, hy, hxi, hyi = g.hx, g.hy, g.hxi, g.hyi
    w = 0.5
    mat_hx = sdiag(hxi); mat_hy = sdiag(hyi)
    perx = BC.u.left == 'per' and BC.u.right == 'per'
    pery = BC.v.low == 'per' and BC.v.up == 'per'
    if not d.order4:
        mat_hx2 = sdiag(np.r_[hx[-1], hx, hx[0]] if perx else np.r_[hx[0], hx, hx[-1]])
        mat_hy2 = sdiag(np.r_[hy[-1], hy, hy[0]] if pery else np.r_[hy[0], hy, hy[-1]])

        def I1(n):
            return mspdiags([w * np.ones(n), w * np.ones(n)], [0, 1], n - 1, n).tolil()

        b = BC_general(g.Nux_t, g.Nux_in, g.Nux_b, BC.u.left, BC.u.right, hx[0], hx[-1])
        I = I1(g.Nux_t).tocsr()
        d.Iu_ux = kron(mat_hy, I @ b.B1D); b.Bbc = kron(mat_hy, I @ b.Btemp); d.Iu_ux_BC = b

        # Iv_uy (with the 2022 half-cell fix at 'pres' boundaries, :82-90)
        I = I1(g.Nvx_t)
        if BC.u.right == 'pres':
            I[I.shape[0] - 1, :] = I[I.shape[0] - 1, :] / 2
        if BC.u.left == 'pres':
            I[0, :] = I[0, :] / 2
        I = g.Bvux @ I.tocsr() @ mat_hx2
        I2D = kron(speye(g.Nuy_t - 1), I)
        Nb = g.Nuy_in + 1 - g.Nvy_in
        lu = BC_general(g.Nuy_in + 1, g.Nvy_in, Nb, BC.v.low, BC.v.up, hy[0], hy[-1])
        lu.B2D = kron(lu.B1D, speye(g.Nvx_in)); lu.Bbc = kron(lu.Btemp, speye(g.Nvx_in))
        lr = BC_general_stag(g.Nvx_t, g.Nvx_in, g.Nvx_b, BC.v.left, BC.v.right, hx[0], hx[-1])
        lr.B2D = I2D @ kron(speye(g.Nuy_t - 1), lr.B1D); lr.Bbc = I2D @ kron(speye(g.Nuy_t - 1), lr.Btemp)
        d.Iv_uy = (lr.B2D @ lu.B2D).tocsr(); d.Iv_uy_BC_lr = lr; d.Iv_uy_BC_lu = lu

        # Iu_vx (:119-150)
        I = I1(g.Nuy_t)
        if BC.v.up == 'pres':
            I[I.shape[0] - 1, :] = I[I.shape[0] - 1, :] / 2
        if BC.v.low == 'pres':
            I[0, :] = I[0, :] / 2
        I = g.Buvy @ I.tocsr() @ mat_hy2
        I2D = kron(I, speye(g.Nvx_t - 1))
        lu = BC_general_stag(g.Nuy_t, g.Nuy_in, g.Nuy_b, BC.u.low, BC.u.up, hy[0], hy[-1])
        lu.B2D = I2D @ kron(lu.B1D, speye(g.Nvx_t - 1)); lu.Bbc = I2D @ kron(lu.Btemp, speye(g.Nvx_t - 1))
        Nb = g.Nvx_in + 1 - g.Nux_in
        lr = BC_general(g.Nvx_in + 1, g.Nux_in, Nb, BC.u.left, BC.u.right, hx[0], hx[-1])
        lr.B2D = kron(speye(g.Nuy_in), lr.B1D); lr.Bbc = kron(speye(g.Nuy_in), lr.Btemp)
        d.Iu_vx = (lu.B2D @ lr.B2D).tocsr(); d.Iu_vx_BC_lr = lr; d.Iu_vx_BC_lu = lu

        b = BC_general(g.Nvy_t, g.Nvy_in, g.Nvy_b, BC.v.low, BC.v.up, hy[0], hy[-1])
        I = I1(g.Nvy_t).tocsr()
        d.Iv_vy = kron(I @ b.B1D, mat_hx); b.Bbc = kron(I @ b.Btemp, mat_hx); d.Iv_vy_BC = b
        return

    # ---- fourth order (:166-337)
    beta = d.beta
    hx3, hy3, hxi3, hyi3 = g.hx3, g.hy3, g.hxi3, g.hyi3
    mat_hx3 = sdiag(hxi3); mat_hy3 = sdiag(hyi3)
    w1 = 0.5 * beta; w2 = 0.5 * (1 - beta)
    if perx:
        mat_hx2 = sdiag(np.r_[hx[-2], hx[-1], hx, hx[0], hx[1]])
        mat_hx4 = sdiag(np.r_[hx3[-2], hx3[-1], hxi3, hx3[0], hx3[1]])
    else:
        mat_hx2 = sdiag(np.r_[hx[1], hx[0], hx, hx[-1], hx[-2]])
        mat_hx4 = sdiag(np.r_[hx[0] + hx[1] + hx[2], 2 * hx[0] + hx[1], hxi3, 2 * hx[-1] + hx[-2], hx[-1] + hx[-2] + hx[-3]])
    if pery:
        mat_hy2 = sdiag(np.r_[hy[-2], hy[-1], hy, hy[0], hy[1]])
        mat_hy4 = sdiag(np.r_[hy3[-2], hy3[-1], hyi3, hy3[0], hy3[1]])
    else:
        mat_hy2 = sdiag(np.r_[hy[1], hy[0], hy, hy[-1], hy[-2]])
        mat_hy4 = sdiag(np.r_[hy[0] + hy[1] + hy[2], 2 * hy[0] + hy[1], hyi3, 2 * hy[-1] + hy[-2], hy[-1] + hy[-2] + hy[-3]])

    def I4(ndiag, m, n):
        return mspdiags([w2 * np.ones(ndiag), w1 * np.ones(ndiag), w1 * np.ones(ndiag), w2 * np.ones(ndiag)], [0, 1, 2, 3], m, n)

    # Iu_ux
    I = I4(g.Nux_t + 1, g.Nux_t - 1, g.Nux_t + 2)
    b = BC_int2(g.Nux_t + 2, g.Nux_in, g.Nux_t + 2 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
    d.Iu_ux = kron(mat_hy, I @ b.B1D); b.Bbc = kron(mat_hy, I @ b.Btemp); d.Iu_ux_BC = b
    I = I4(g.Nux_t + 4, g.Nux_in + 3, g.Nux_t + 4)
    b = BC_int3(g.Nux_t + 4, g.Nux_in, g.Nux_t + 4 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
    d.Iu_ux3 = kron(mat_hy3, I @ b.B1D); b.Bbc = kron(mat_hy3, I @ b.Btemp); d.Iu_ux_BC3 = b
    # Iv_uy
    I = g.Bvux @ I4(g.Nvx_t, g.Nvx_t - 1, g.Nvx_t + 2) @ mat_hx2
    I2D = kron(speye(g.Nuy_t - 1), I)
    Nb = g.Nuy_in + 1 - g.Nvy_in
    lu = BC_general(g.Nuy_in + 1, g.Nvy_in, Nb, BC.v.low, BC.v.up, hy[0], hy[-1])
    lu.B2D = kron(lu.B1D, speye(g.Nvx_in)); lu.Bbc = kron(lu.Btemp, speye(g.Nvx_in))
    lr = BC_int_mixed_stag2(g.Nvx_t + 2, g.Nvx_in, g.Nvx_t + 2 - g.Nvx_in, BC.v.left, BC.v.right, hx[0], hx[-1])
    lr.B2D = I2D @ kron(speye(g.Nuy_t - 1), lr.B1D); lr.Bbc = I2D @ kron(speye(g.Nuy_t - 1), lr.Btemp)
    d.Iv_uy = (lr.B2D @ lu.B2D).tocsr(); d.Iv_uy_BC_lr = lr; d.Iv_uy_BC_lu = lu
    # Iv_uy3
    I = g.Bvux @ I4(g.Nvx_t, g.Nvx_t - 1, g.Nvx_t + 2) @ mat_hx4
    I2D = kron(speye(g.Nuy_t + 1), I)
    Nb = g.Nuy_in + 3 - g.Nvy_in
    lu = BC_int_mixed2(g.Nuy_in + 3, g.Nvy_in, Nb, BC.v.low, BC.v.up, hy[0], hy[-1])
    lu.B2D = kron(lu.B1D, speye(g.Nvx_in)); lu.Bbc = kron(lu.Btemp, speye(g.Nvx_in))
    lr = BC_int_mixed_stag3(g.Nvx_t + 2, g.Nvx_in, g.Nvx_t + 2 - g.Nvx_in, BC.v.left, BC.v.right, hx[0], hx[-1])
    lr.B2D = I2D @ kron(speye(g.Nuy_t + 1), lr.B1D); lr.Bbc = I2D @ kron(speye(g.Nuy_t + 1), lr.Btemp)
    d.Iv_uy3 = (lr.B2D @ lu.B2D).tocsr(); d.Iv_uy_BC_lr3 = lr; d.Iv_uy_BC_lu3 = lu
    # Iu_vx
    I = g.Buvy @ I4(g.Nuy_t, g.Nuy_t - 1, g.Nuy_t + 2) @ mat_hy2
    I2D = kron(I, speye(g.Nvx_t - 1))
    lu = BC_int_mixed_stag2(g.Nuy_t + 2, g.Nuy_in, g.Nuy_t + 2 - g.Nuy_in, BC.u.low, BC.u.up, hy[0], hy[-1])
    lu.B2D = I2D @ kron(lu.B1D, speye(g.Nvx_t - 1)); lu.Bbc = I2D @ kron(lu.Btemp, speye(g.Nvx_t - 1))
    Nb = g.Nvx_in + 1 - g.Nux_in
    lr = BC_general(g.Nvx_in + 1, g.Nux_in, Nb, BC.u.left, BC.u.right, hx[0], hx[-1])
    lr.B2D = kron(speye(g.Nuy_in), lr.B1D); lr.Bbc = kron(speye(g.Nuy_in), lr.Btemp)
    d.Iu_vx = (lu.B2D @ lr.B2D).tocsr(); d.Iu_vx_BC_lr = lr; d.Iu_vx_BC_lu = lu
    # Iu_vx3
    I = g.Buvy @ I4(g.Nuy_t, g.Nuy_t - 1, g.Nuy_t + 2) @ mat_hy4
    I2D = kron(I, speye(g.Nvx_t + 1))
    lu = BC_int_mixed_stag3(g.Nuy_t + 2, g.Nuy_in, g.Nuy_t + 2 - g.Nuy_in, BC.u.low, BC.u.up, hy[0], hy[-1])
    lu.B2D = I2D @ kron(lu.B1D, speye(g.Nvx_t + 1)); lu.Bbc = I2D @ kron(lu.Btemp, speye(g.Nvx_t + 1))
    Nb = g.Nvx_in + 3 - g.Nux_in
    lr = BC_int_mixed2(g.Nvx_in + 3, g.Nux_in, Nb, BC.u.left, BC.u.right, hx[0], hx[-1])
    lr.B2D = kron(speye(g.Nuy_in), lr.B1D); lr.Bbc = kron(speye(g.Nuy_in), lr.Btemp)
    d.Iu_vx3 = (lu.B2D @ lr.B2D).tocsr(); d.Iu_vx_BC_lr3 = lr; d.Iu_vx_BC_lu3 = lu
    # Iv_vy
    I = I4(g.Nvy_t + 1, g.Nvy_t - 1, g.Nvy_t + 2)
    b = BC_int2(g.Nvy_t + 2, g.Nvy_in, g.Nvy_t + 2 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
    d.Iv_vy = kron(I @ b.B1D, mat_hx); b.Bbc = kron(I @ b.Btemp, mat_hx); d.Iv_vy_BC = b
    I = I4(g.Nvy_t + 4, g.Nvy_in + 3, g.Nvy_t + 4)
    b = BC_int3(g.Nvy_t + 4, g.Nvy_in, g.Nvy_t + 4 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
    d.Iv_vy3 = kron(I @ b.B1D, mat_hx3); b.Bbc = kron(I @ b.Btemp, mat_hx3); d.Iv_vy_BC3 = b


def operator_divergence(opts):
    """spatial/operators/operator_divergence.m:1-291 (M, G=-M^T, A=M Om^-1 G,
    and the per-mode Poisson setup)."""
    BC = opts.BC; g = opts.grid; d = opts.discretization
    hx, hy = g.hx, g.hy
    mat_hx = sdiag(hx); mat_hy = sdiag(hy)

    def dpos(lo, hi):
        pos = 0
        if hi == 'pres' and lo == 'pres': pos = 1
        if hi != 'pres' and lo == 'pres': pos = 1
        if hi == 'pres' and lo != 'pres': pos = 0
        if hi == 'per' and lo == 'per': pos = 1
        return pos

    M1D = mspdiags([-np.ones(g.Nux_t), np.ones(g.Nux_t)], [0, 1], g.Nux_t - 1, g.Nux_t)
    BMx = mspdiags([np.ones(g.Npx)], [dpos(BC.u.left, BC.u.right)], g.Npx, g.Nux_t - 1)
    M1D = BMx @ M1D
    b = BC_general(g.Nux_t, g.Nux_in, g.Nux_b, BC.u.left, BC.u.right, hx[0], hx[-1])
    b.Bbc = kron(mat_hy, M1D @ b.Btemp); d.Mx_BC = b
    Mx = kron(mat_hy, M1D @ b.B1D)
    if d.order4:
        M1D3 = BMx @ mspdiags([-np.ones(g.Nux_t + 1), np.ones(g.Nux_t + 1)], [0, 3], g.Nux_t - 1, g.Nux_t + 2)
        b3 = BC_div2(g.Nux_t + 2, g.Nux_in, g.Nux_t + 2 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
        Mx3 = kron(sdiag(g.hyi3), M1D3 @ b3.B1D); b3.Bbc = kron(sdiag(g.hyi3), M1D3 @ b3.Btemp); d.Mx_BC3 = b3

    M1D = mspdiags([-np.ones(g.Nvy_t), np.ones(g.Nvy_t)], [0, 1], g.Nvy_t - 1, g.Nvy_t)
    BMy = mspdiags([np.ones(g.Npy)], [dpos(BC.v.low, BC.v.up)], g.Npy, g.Nvy_t - 1)
    M1D = BMy @ M1D
    b = BC_general(g.Nvy_t, g.Nvy_in, g.Nvy_b, BC.v.low, BC.v.up, hy[0], hy[-1])
    b.Bbc = kron(M1D @ b.Btemp, mat_hx); d.My_BC = b
    My = kron(M1D @ b.B1D, mat_hx)
    if d.order4:
        M1D3 = BMy @ mspdiags([-np.ones(g.Nvy_t + 1), np.ones(g.Nvy_t + 1)], [0, 3], g.Nvy_t - 1, g.Nvy_t + 2)
        b3 = BC_div2(g.Nvy_t + 2, g.Nvy_in, g.Nvy_t + 2 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
        My3 = kron(M1D3 @ b3.B1D, sdiag(g.hxi3)); b3.Bbc = kron(M1D3 @ b3.Btemp, sdiag(g.hxi3)); d.My_BC3 = b3
        Mx = d.alfa * Mx - Mx3
        My = d.alfa * My - My3
    d.Mx, d.My = Mx.tocsr(), My.tocsr()
    d.M = sp.hstack([Mx, My], format='csr')
    d.Gx, d.Gy = (-Mx.T).tocsr(), (-My.T).tocsr()
    d.G = sp.vstack([d.Gx, d.Gy], format='csr')
    d.A = (d.M @ sdiag(g.Om_inv) @ d.G).tocsr()
    s = opts.solversettings
    if s.poisson == 6:
        if not (BC.v.low == 'per' and BC.v.up == 'per' and BC.u.left == 'per'):
            raise ValueError('Fourier transform for pressure Poisson only implemented for periodic boundary conditions')
        if max(abs(np.diff(hx))) > 1e-14 or max(abs(np.diff(hy))) > 1e-14:
            raise ValueError('grid needs to be uniform to use Fourier transform for pressure matrix')
        dx, dy = hx[0], hy[0]
        I, J = np.meshgrid(np.arange(g.Npx), np.arange(g.Npy), indexing='ij')
        Ahat = (dx * dy) * 4 * ((1 / dx ** 2) * np.sin(I * np.pi / g.Npx) ** 2 + (1 / dy ** 2) * np.sin(J * np.pi / g.Npy) ** 2)
        if int(getattr(opts, 'prm', {}).get('fft_symbol', 0)) == 1:
            # A is block-circulant on a periodic uniform grid: its eigenvalues
            # are the 2-D DFT of its first column (exact for either order;
            # the reference itself always uses the second-order closed form)
            col = np.asarray(d.A[:, 0].todense()).ravel().reshape(g.Npx, g.Npy, order='F')
            Ahat = -np.real(np.fft.fft2(col))
        Ahat[0, 0] = 1
        s.Ahat = Ahat
    if s.poisson == 1:
        s.decomp = _singular_direct(d.A)
    if s.poisson in (3, 4):
        s.A_pc = make_cholinc_diag(d.A)
    if s.poisson == 2:
        # operator_divergence.m:243-247
        s.L, s.U = ilu0(d.A)


def _singular_direct(A):
    """Direct solve for A that may be singular with a constant null vector
    (all-wall cavities).  The reference's ``decomposition(A)\\f`` picks an
    arbitrary constant; this returns the zero-mean solution of the bordered
    system instead (pressure is compared modulo its constant)."""
    N = A.shape[0]
    rowsum = np.abs(A @ np.ones(N)).max()
    if rowsum > 1e-10 * abs(A.diagonal()).max():
        lu = spla.splu(A.tocsc())
        return lambda f: lu.solve(f)
    e = sp.csr_matrix(np.ones((N, 1)))
    K = sp.bmat([[A, e], [e.T, None]], format='csc')
    lu = spla.splu(K)
    return lambda f: lu.solve(np.r_[f, 0.0])[:N]


def operator_convection_diffusion(opts):
    """spatial/operators/operator_convection_diffusion.m:1-470 (laminar)."""
    BC = opts.BC; g = opts.grid; d = opts.discretization
    Re = opts.fluid.Re
    Nx, Ny = g.Nx, g.Ny
    hx, hy = g.hx, g.hy

    def D1(nr, nc, offs=(0, 1)):
        return mspdiags([-np.ones(nc + 4), np.ones(nc + 4)], list(offs), nr, nc)

    d.Cux = kron(speye(g.Nuy_in), D1(g.Nux_t - 2, g.Nux_t - 1))
    d.Cuy = kron(D1(g.Nuy_t - 2, g.Nuy_t - 1), speye(g.Nux_in))
    d.Cvx = kron(speye(g.Nvy_in), D1(g.Nvx_t - 2, g.Nvx_t - 1))
    d.Cvy = kron(D1(g.Nvy_t - 2, g.Nvy_t - 1), speye(g.Nvx_in))
    if not d.order4:
        d.Dux = kron(sdiag(g.hyi), D1(g.Nux_t - 2, g.Nux_t - 1))
        d.Duy = kron(D1(g.Nuy_t - 2, g.Nuy_t - 1), sdiag(g.gxi))
        d.Dvx = kron(sdiag(g.gyi), D1(g.Nvx_t - 2, g.Nvx_t - 1))
        d.Dvy = kron(D1(g.Nvy_t - 2, g.Nvy_t - 1), sdiag(g.hxi))

        def S1(hinv, nr, nc):
            return mspdiags([-hinv, hinv], [0, 1], nr, nc)

        b = BC_general(g.Nux_t, g.Nux_in, g.Nux_b, BC.u.left, BC.u.right, hx[0], hx[-1])
        S = S1(1 / g.hxd, g.Nux_t - 1, g.Nux_t)
        d.Su_ux = kron(speye(Ny), S @ b.B1D); b.Bbc = kron(speye(Ny), S @ b.Btemp); d.Su_ux_BC = b
        b = BC_diff_stag(g.Nuy_t, g.Nuy_in, g.Nuy_b, BC.u.low, BC.u.up, hy[0], hy[-1])
        S = S1(1 / g.gyd, g.Nuy_t - 1, g.Nuy_t)
        d.Su_uy = kron(S @ b.B1D, speye(g.Nux_in)); b.Bbc = kron(S @ b.Btemp, speye(g.Nux_in)); d.Su_uy_BC = b
        b = BC_diff_stag(g.Nvx_t, g.Nvx_in, g.Nvx_b, BC.v.left, BC.v.right, hx[0], hx[-1])
        S = S1(1 / g.gxd, g.Nvx_t - 1, g.Nvx_t)
        d.Sv_vx = kron(speye(g.Nvy_in), S @ b.B1D); b.Bbc = kron(speye(g.Nvy_in), S @ b.Btemp); d.Sv_vx_BC = b
        b = BC_general(g.Nvy_t, g.Nvy_in, g.Nvy_b, BC.v.low, BC.v.up, hy[0], hy[-1])
        S = S1(1 / g.hyd, g.Nvy_t - 1, g.Nvy_t)
        d.Sv_vy = kron(S @ b.B1D, speye(Nx)); b.Bbc = kron(S @ b.Btemp, speye(Nx)); d.Sv_vy_BC = b
        d.Diffu = (d.Dux @ ((1 / Re) * d.Su_ux) + d.Duy @ ((1 / Re) * d.Su_uy)).tocsr()
        d.Diffv = (d.Dvx @ ((1 / Re) * d.Sv_vx) + d.Dvy @ ((1 / Re) * d.Sv_vy)).tocsr()
        return

    # ---- fourth order (:257-449, :462-470)
    alfa = d.alfa
    d.Dux = kron(sdiag(g.hyi), D1(g.Nux_t - 2, g.Nux_t + 1, (1, 2)))
    D13 = D1(g.Nux_t - 2, g.Nux_t + 1, (0, 3))
    d.Cux3 = kron(speye(Ny), D13); d.Dux3 = kron(sdiag(g.hyi3), D13)
    d.Duy = kron(D1(g.Nuy_t - 2, g.Nuy_t + 1, (1, 2)), sdiag(g.gxi))
    D13 = D1(g.Nuy_t - 2, g.Nuy_t + 1, (0, 3)).tolil()
    if BC.u.low == 'dir':
        D13[0, 0] = 1; D13[0, 1] = -2
    if BC.u.up == 'dir':
        D13[D13.shape[0] - 1, D13.shape[1] - 2] = 2; D13[D13.shape[0] - 1, D13.shape[1] - 1] = -1
    D13 = D13.tocsr()
    d.Cuy3 = kron(D13, speye(g.Nux_in)); d.Duy3 = kron(D13, sdiag(g.gxi3))
    d.Dvx = kron(sdiag(g.gyi), D1(g.Nvx_t - 2, g.Nvx_t + 1, (1, 2)))
    D13 = D1(g.Nvx_t - 2, g.Nvx_t + 1, (0, 3)).tolil()
    if BC.v.left == 'dir':
        D13[0, 0] = 1; D13[0, 1] = -2
    if BC.v.right == 'dir':
        D13[D13.shape[0] - 1, D13.shape[1] - 2] = 2; D13[D13.shape[0] - 1, D13.shape[1] - 1] = -1
    D13 = D13.tocsr()
    d.Cvx3 = kron(speye(g.Nvy_in), D13); d.Dvx3 = kron(sdiag(g.gyi3), D13)
    d.Dvy = kron(D1(g.Nvy_t - 2, g.Nvy_t + 1, (1, 2)), sdiag(g.hxi))
    D13 = D1(g.Nvy_t - 2, g.Nvy_t + 1, (0, 3))
    d.Cvy3 = kron(D13, speye(g.Nvx_in)); d.Dvy3 = kron(D13, sdiag(g.hxi3))

    def S(hinv, offs, nr, nc):
        return mspdiags([-hinv, hinv], list(offs), nr, nc)

    def dg(v):
        return sdiag(v)
    b = BC_diff3(g.Nux_t + 4, g.Nux_in, g.Nux_t + 4 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
    S1 = S(1 / g.hxd13, (1, 2), g.Nux_in + 3, g.Nux_t + 4)
    d.Su_ux = dg(g.Omux1) @ kron(speye(Ny), S1 @ b.B1D); b.Bbc = dg(g.Omux1) @ kron(speye(Ny), S1 @ b.Btemp); d.Su_ux_BC = b
    b3 = BC_diff3(g.Nux_t + 4, g.Nux_in, g.Nux_t + 4 - g.Nux_in, BC.u.left, BC.u.right, hx[0], hx[-1])
    S3 = S(1 / g.hxd3, (0, 3), g.Nux_in + 3, g.Nux_t + 4)
    d.Su_ux3 = dg(g.Omux3) @ kron(speye(g.Nuy_in), S3 @ b3.B1D); b3.Bbc = dg(g.Omux3) @ kron(speye(g.Nuy_in), S3 @ b3.Btemp); d.Su_ux_BC3 = b3

    b = BC_diff_stag3(g.Nuy_t + 4, g.Nuy_in, g.Nuy_t + 4 - g.Nuy_in, BC.u.low, BC.u.up, hy[0], hy[-1])
    S1 = S(1 / g.gyd13, (1, 2), g.Nuy_in + 3, g.Nuy_t + 4)
    d.Su_uy = dg(g.Omuy1) @ kron(S1 @ b.B1D, speye(g.Nux_in)); b.Bbc = dg(g.Omuy1) @ kron(S1 @ b.Btemp, speye(g.Nux_in)); d.Su_uy_BC = b
    b3 = BC_diff_stag3(g.Nuy_t + 4, g.Nuy_in, g.Nuy_t + 4 - g.Nuy_in, BC.u.low, BC.u.up, hy[0], hy[-1])
    S3 = S(1 / g.gyd3, (0, 3), g.Nuy_in + 3, g.Nuy_t + 4)
    d.Su_uy3 = dg(g.Omuy3) @ kron(S3 @ b3.B1D, speye(g.Nux_in)); b3.Bbc = dg(g.Omuy3) @ kron(S3 @ b3.Btemp, speye(g.Nux_in)); d.Su_uy_BC3 = b3

    b = BC_diff_stag3(g.Nvx_t + 4, g.Nvx_in, g.Nvx_t + 4 - g.Nvx_in, BC.v.left, BC.v.right, hx[0], hx[-1])
    S1 = S(1 / g.gxd13, (1, 2), g.Nvx_in + 3, g.Nvx_t + 4)
    d.Sv_vx = dg(g.Omvx1) @ kron(speye(g.Nvy_in), S1 @ b.B1D); b.Bbc = dg(g.Omvx1) @ kron(speye(g.Nvy_in), S1 @ b.Btemp); d.Sv_vx_BC = b
    b3 = BC_diff_stag3(g.Nvx_t + 4, g.Nvx_in, g.Nvx_t + 4 - g.Nvx_in, BC.v.left, BC.v.right, hx[0], hx[-1])
    S3 = S(1 / g.gxd3, (0, 3), g.Nvx_in + 3, g.Nvx_t + 4)
    d.Sv_vx3 = dg(g.Omvx3) @ kron(speye(g.Nvy_in), S3 @ b3.B1D); b3.Bbc = dg(g.Omvx3) @ kron(speye(g.Nvy_in), S3 @ b3.Btemp); d.Sv_vx_BC3 = b3

    b = BC_diff3(g.Nvy_t + 4, g.Nvy_in, g.Nvy_t + 4 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
    S1 = S(1 / g.hyd13, (1, 2), g.Nvy_in + 3, g.Nvy_t + 4)
    d.Sv_vy = dg(g.Omvy1) @ kron(S1 @ b.B1D, speye(g.Nvx_in)); b.Bbc = dg(g.Omvy1) @ kron(S1 @ b.Btemp, speye(g.Nvx_in)); d.Sv_vy_BC = b
    b3 = BC_diff3(g.Nvy_t + 4, g.Nvy_in, g.Nvy_t + 4 - g.Nvy_in, BC.v.low, BC.v.up, hy[0], hy[-1])
    S3 = S(1 / g.hyd3, (0, 3), g.Nvy_in + 3, g.Nvy_t + 4)
    d.Sv_vy3 = dg(g.Omvy3) @ kron(S3 @ b3.B1D, speye(g.Nvx_in)); b3.Bbc = dg(g.Omvy3) @ kron(S3 @ b3.Btemp, speye(g.Nvx_in)); d.Sv_vy_BC3 = b3

    d.Diffux_div = (alfa * d.Dux - d.Dux3) @ dg(1 / g.Omux)
    d.Diffuy_div = (alfa * d.Duy - d.Duy3) @ dg(1 / g.Omuy)
    d.Diffvx_div = (alfa * d.Dvx - d.Dvx3) @ dg(1 / g.Omvx)
    d.Diffvy_div = (alfa * d.Dvy - d.Dvy3) @ dg(1 / g.Omvy)
    d.Diffu = ((1 / Re) * d.Diffux_div @ (alfa * d.Su_ux - d.Su_ux3) + (1 / Re) * d.Diffuy_div @ (alfa * d.Su_uy - d.Su_uy3)).tocsr()
    d.Diffv = ((1 / Re) * d.Diffvx_div @ (alfa * d.Sv_vx - d.Sv_vx3) + (1 / Re) * d.Diffvy_div @ (alfa * d.Sv_vy - d.Sv_vy3)).tocsr()


def operators(opts):
    """spatial/operators/operators.m:2-19 (laminar branch)."""
    operator_mesh(opts)
    operator_averaging(opts)
    operator_interpolation(opts)
    operator_divergence(opts)
    operator_convection_diffusion(opts)


# ----------------------------------------------------------------------------
# boundary-condition vectors (spatial/boundaryconditions/set_bc_vectors.m)
# ----------------------------------------------------------------------------


def set_bc_vectors(t, opts):
    """spatial/boundaryconditions/set_bc_vectors.m:1-546 (laminar)."""
    BC = opts.BC; g = opts.grid; d = opts.discretization; case = opts.case_obj
    Re = opts.fluid.Re
    x, y, xin, yin, xp, yp, hx, hy = g.x, g.y, g.xin, g.yin, g.xp, g.yp, g.hx, g.hy
    uBC, vBC = case.uBC, case.vBC
    uLo = uBC(x, y[0:1], t, opts); uUp = uBC(x, y[-1:], t, opts)
    uLo_i = uBC(xin, y[0:1], t, opts); uUp_i = uBC(xin, y[-1:], t, opts)
    uLe_i = uBC(x[0:1], yp, t, opts); uRi_i = uBC(x[-1:], yp, t, opts)
    vLe = vBC(x[0:1], y, t, opts); vRi = vBC(x[-1:], y, t, opts)
    vLo_i = vBC(xp, y[0:1], t, opts); vUp_i = vBC(xp, y[-1:], t, opts)
    vLe_i = vBC(x[0:1], yin, t, opts); vRi_i = vBC(x[-1:], yin, t, opts)
    k = np.kron
    alfa = d.alfa

    b = d.Mx_BC
    yMx = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
    b = d.My_BC
    yMy = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))
    if d.order4:
        b = d.Mx_BC3
        yMx = alfa * yMx - b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
        b = d.My_BC3
        yMy = alfa * yMy - b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))
    d.yM = yMx + yMy
    # time derivative of the divergence's boundary vector (:120-146): the same
    # construction on the dudtBC / dvdtBC samples (unsteady BC only)
    if BC.BC_unsteady:
        dudtLe_i = case.dudtBC(x[0:1], yp, t, opts); dudtRi_i = case.dudtBC(x[-1:], yp, t, opts)
        dvdtLo_i = case.dvdtBC(xp, y[0:1], t, opts); dvdtUp_i = case.dvdtBC(xp, y[-1:], t, opts)
        b = d.Mx_BC
        ydMx = b.Bbc @ (k(dudtLe_i, b.ybc1) + k(dudtRi_i, b.ybc2))
        b = d.My_BC
        ydMy = b.Bbc @ (k(b.ybc1, dvdtLo_i) + k(b.ybc2, dvdtUp_i))
        if d.order4:
            b = d.Mx_BC3
            ydMx = alfa * ydMx - b.Bbc @ (k(dudtLe_i, b.ybc1) + k(dudtRi_i, b.ybc2))
            b = d.My_BC3
            ydMy = alfa * ydMy - b.Bbc @ (k(b.ybc1, dvdtLo_i) + k(b.ybc2, dvdtUp_i))
        d.ydM = ydMx + ydMy
    else:
        d.ydM = np.zeros(g.Np)

    # pressure (:156-180), p_inf = 0 (boundary_conditions.m:107)
    pLe = pRi = pLo = pUp = opts.BC.p_inf
    y1 = np.zeros(g.Nux_in); y2 = np.zeros(g.Nux_in)
    if BC.u.left == 'pres': y1[0] = -1
    if BC.u.right == 'pres': y2[-1] = 1
    d.y_px = k(hy * pLe, y1) + k(hy * pRi, y2)
    y1 = np.zeros(g.Nvy_in); y2 = np.zeros(g.Nvy_in)
    if BC.v.low == 'pres': y1[0] = -1
    if BC.v.up == 'pres': y2[-1] = 1
    d.y_py = k(y1, hx * pLo) + k(y2, hx * pUp)

    # averaging (:196-230)
    b = d.Au_ux_BC; d.yAu_ux = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
    b = d.Au_uy_BC; d.yAu_uy = b.Bbc @ (k(b.ybc1, uLo_i) + k(b.ybc2, uUp_i))
    b = d.Av_vx_BC; d.yAv_vx = b.Bbc @ (k(vLe_i, b.ybc1) + k(vRi_i, b.ybc2))
    b = d.Av_vy_BC; d.yAv_vy = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))
    if d.order4:
        b = d.Au_ux_BC3; d.yAu_ux3 = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
        b = d.Au_uy_BC3; d.yAu_uy3 = b.Bbc @ (k(b.ybc1, uLo_i) + k(b.ybc2, uUp_i))
        b = d.Av_vx_BC3; d.yAv_vx3 = b.Bbc @ (k(vLe_i, b.ybc1) + k(vRi_i, b.ybc2))
        b = d.Av_vy_BC3; d.yAv_vy3 = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))

    # diffusion (:258-411)
    if not d.order4:
        b = d.Su_ux_BC; ySu_ux = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
        b = d.Su_uy_BC; ySu_uy = b.Bbc @ (k(b.ybc1, uLo_i) + k(b.ybc2, uUp_i))
        b = d.Sv_vx_BC; ySv_vx = b.Bbc @ (k(vLe_i, b.ybc1) + k(vRi_i, b.ybc2))
        b = d.Sv_vy_BC; ySv_vy = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))
        d.yDiffu = d.Dux @ ((1 / Re) * ySu_ux) + d.Duy @ ((1 / Re) * ySu_uy)
        d.yDiffv = d.Dvx @ ((1 / Re) * ySv_vx) + d.Dvy @ ((1 / Re) * ySv_vy)
    else:
        def pair(b1, b3, lo, hi, first):
            y1 = (k(lo, b1.ybc1) + k(hi, b1.ybc2)) if first else (k(b1.ybc1, lo) + k(b1.ybc2, hi))
            y3 = (k(lo, b3.ybc1) + k(hi, b3.ybc2)) if first else (k(b3.ybc1, lo) + k(b3.ybc2, hi))
            return alfa * (b1.Bbc @ y1) - b3.Bbc @ y3
        ySu_ux = pair(d.Su_ux_BC, d.Su_ux_BC3, uLe_i, uRi_i, True)
        ySu_uy = pair(d.Su_uy_BC, d.Su_uy_BC3, uLo_i, uUp_i, False)
        ySv_vx = pair(d.Sv_vx_BC, d.Sv_vx_BC3, vLe_i, vRi_i, True)
        ySv_vy = pair(d.Sv_vy_BC, d.Sv_vy_BC3, vLo_i, vUp_i, False)
        d.yDiffu = (1 / Re) * (d.Diffux_div @ ySu_ux + d.Diffuy_div @ ySu_uy)
        d.yDiffv = (1 / Re) * (d.Diffvx_div @ ySv_vx + d.Diffvy_div @ ySv_vy)

    # interpolation (:433-534)
    b = d.Iu_ux_BC; d.yIu_ux = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
    lr, lu = d.Iv_uy_BC_lr, d.Iv_uy_BC_lu
    d.yIv_uy = lr.Bbc @ (k(vLe, lr.ybc1) + k(vRi, lr.ybc2)) + lr.B2D @ (lu.Bbc @ (k(lu.ybc1, vLo_i) + k(lu.ybc2, vUp_i)))
    lr, lu = d.Iu_vx_BC_lr, d.Iu_vx_BC_lu
    d.yIu_vx = lu.Bbc @ (k(lu.ybc1, uLo) + k(lu.ybc2, uUp)) + lu.B2D @ (lr.Bbc @ (k(uLe_i, lr.ybc1) + k(uRi_i, lr.ybc2)))
    b = d.Iv_vy_BC; d.yIv_vy = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))
    if d.order4:
        b = d.Iu_ux_BC3; d.yIu_ux3 = b.Bbc @ (k(uLe_i, b.ybc1) + k(uRi_i, b.ybc2))
        # ghost extension of vLe/vRi (:455-474)
        def ext(lo_bc, hi_bc, a, bb):
            if lo_bc == 'dir':
                ae = [2 * a[0] - a[1]]; be = [2 * bb[0] - bb[1]]
            elif lo_bc == 'per':
                ae = [0.0]; be = [0.0]
            else:
                ae = [a[1]]; be = [bb[1]]
            ae = np.r_[ae, a]; be = np.r_[be, bb]
            if hi_bc == 'dir':
                ae = np.r_[ae, 2 * a[-1] - a[-2]]; be = np.r_[be, 2 * bb[0] - bb[1]]
            elif hi_bc == 'per':
                ae = np.r_[ae, 0.0]; be = np.r_[be, 0.0]
            else:
                ae = np.r_[ae, a[-2]]; be = np.r_[be, bb[-2]]
            return ae, be
        vLe_ext, vRi_ext = ext(BC.v.low, BC.v.up, vLe, vRi)
        lr, lu = d.Iv_uy_BC_lr3, d.Iv_uy_BC_lu3
        d.yIv_uy3 = lr.Bbc @ (k(vLe_ext, lr.ybc1) + k(vRi_ext, lr.ybc2)) + lr.B2D @ (lu.Bbc @ (k(lu.ybc1, vLo_i) + k(lu.ybc2, vUp_i)))
        uLo_ext, uUp_ext = ext(BC.u.left, BC.u.right, uLo, uUp)
        lr, lu = d.Iu_vx_BC_lr3, d.Iu_vx_BC_lu3
        d.yIu_vx3 = lu.Bbc @ (k(lu.ybc1, uLo_ext) + k(lu.ybc2, uUp_ext)) + lu.B2D @ (lr.Bbc @ (k(uLe_i, lr.ybc1) + k(uRi_i, lr.ybc2)))
        b = d.Iv_vy_BC3; d.yIv_vy3 = b.Bbc @ (k(b.ybc1, vLo_i) + k(b.ybc2, vUp_i))


# ----------------------------------------------------------------------------
# right-hand side (spatial/F.m, convection.m, diffusion.m)
# ----------------------------------------------------------------------------


def convection_components(V, C, opts, order3=False):
    """spatial/convection.m:124-215 (regularize=0)."""
    d = opts.discretization; g = opts.grid
    s = '3' if order3 else ''
    uh, vh = V[:g.Nu], V[g.Nu:]
    cu, cv = C[:g.Nu], C[g.Nu:]
    Cux = d.Cux3 if order3 else d.Cux
    Cuy = d.Cuy3 if order3 else d.Cuy
    Cvx = d.Cvx3 if order3 else d.Cvx
    Cvy = d.Cvy3 if order3 else d.Cvy
    u_ux = getattr(d, 'Au_ux' + s) @ uh + getattr(d, 'yAu_ux' + s)
    uf_ux = getattr(d, 'Iu_ux' + s) @ cu + getattr(d, 'yIu_ux' + s)
    u_uy = getattr(d, 'Au_uy' + s) @ uh + getattr(d, 'yAu_uy' + s)
    vf_uy = getattr(d, 'Iv_uy' + s) @ cv + getattr(d, 'yIv_uy' + s)
    v_vx = getattr(d, 'Av_vx' + s) @ vh + getattr(d, 'yAv_vx' + s)
    uf_vx = getattr(d, 'Iu_vx' + s) @ cu + getattr(d, 'yIu_vx' + s)
    v_vy = getattr(d, 'Av_vy' + s) @ vh + getattr(d, 'yAv_vy' + s)
    vf_vy = getattr(d, 'Iv_vy' + s) @ cv + getattr(d, 'yIv_vy' + s)
    convu = Cux @ (uf_ux * u_ux) + Cuy @ (vf_uy * u_uy)
    convv = Cvx @ (uf_vx * v_vx) + Cvy @ (vf_vy * v_vy)
    return convu, convv


def convection(V, C, t, opts):
    """spatial/convection.m:30-44."""
    cu, cv = convection_components(V, C, opts)
    if opts.discretization.order4:
        cu3, cv3 = convection_components(V, C, opts, order3=True)
        a = opts.discretization.alfa
        cu = a * cu - cu3
        cv = a * cv - cv3
    return cu, cv


def diffusion(V, t, opts):
    """spatial/diffusion.m:22-30 (laminar)."""
    d = opts.discretization; g = opts.grid
    return d.Diffu @ V[:g.Nu] + d.yDiffu, d.Diffv @ V[g.Nu:] + d.yDiffv


def F(V, C, p, t, opts, nopressure=False):
    """spatial/F.m:1-117 -> (maxres, F); unsteady BC: the boundary vectors at
    t first (F.m:27-29)."""
    if opts.BC.BC_unsteady:
        set_bc_vectors(t, opts)
    d = opts.discretization
    convu, convv = convection(V, C, t, opts)
    d2u, d2v = diffusion(V, t, opts)
    Fu = -convu + d2u
    Fv = -convv + d2v
    if not nopressure:
        Fu = Fu - (d.Gx @ p + d.y_px)
        Fv = Fv - (d.Gy @ p + d.y_py)
    Fr = np.r_[Fu, Fv]
    return np.max(np.abs(Fr)), Fr


# ----------------------------------------------------------------------------
# pressure (solvers/pressure/*.m)
# ----------------------------------------------------------------------------


def make_cholinc_diag(A):
    """IC(0) diagonal for the symmetric diagonal storage of A
    (solvers/pressure/make_cholinc.m:1-51, cg.F:394-430)."""
    B, dia = spdiags_lower(A)
    N = A.shape[0]
    dd = list(dia) + [N]
    nd = len(dd)
    D = np.zeros(N)
    D[0] = B[0, 0]
    for j in range(1, nd - 1):
        for i in range(dd[j], dd[j + 1]):
            s = 0.0
            for kk in range(1, j + 1):
                s -= B[i - dd[kk], kk] ** 2 / D[i - dd[kk]]
            D[i] = B[i, 0] + s
    return D


def spdiags_lower(A):
    """[B,d]=spdiags(A); keep the main and the lower diagonals, re-ordered as
    cg_matlab.m:30-34 does: B[:, k] holds A(i+d_k, i) = A(i, i+d_k)."""
    A = sp.csr_matrix(A)
    N = A.shape[0]
    coo = A.tocoo()
    offs = np.unique(coo.col - coo.row)
    dia = np.sort(offs[offs >= 0])
    B = np.zeros((N, len(dia)))
    Ad = A.todok() if N < 5000 else None
    for kk, dd in enumerate(dia):
        i = np.arange(N - dd)
        B[i, kk] = np.asarray(A[i + dd, i]).ravel() if Ad is None else np.array([Ad.get((r + dd, r), 0.0) for r in i])
    return B, dia


def cg_matlab(A, y, tol, x_in, maxiter, D=None):
    """IC(0)-preconditioned CG in diagonal storage (solvers/pressure/
    cg_matlab.m:7-104; the compiled twin is cg.F:158-280).

    Stops when ||r||^2/||y||^2 <= tol^2 (cg.F:247).  Returns x, iter, norm1.
    """
    N = A.shape[0]
    Asym = sp.csr_matrix(A)
    L = sp.tril(Asym, -1, format='csr')
    if D is None:
        D = make_cholinc_diag(A)
    DL = (sp.diags(D) + L).tocsr()
    DU = (sp.diags(D) + L.T).tocsr()
    Dinv = 1.0 / D

    def precond(r):
        q = spla.spsolve_triangular(DL, r, lower=True)
        # (I + D^-1 L^T) x = q
        return spla.spsolve_triangular((sp.identity(N) + sp.diags(Dinv) @ L.T).tocsr(), q, lower=False)

    x = x_in.copy()
    r = y - Asym @ x
    z = precond(r)
    p = z.copy()
    nr0 = float(y @ y)
    if nr0 <= tol * tol:
        return x, 0, 0.0
    rz = float(r @ z)
    it = 0
    while True:
        it += 1
        if it > maxiter:
            raise RuntimeError(f"CG didn't converge in {maxiter} iterations")
        q = Asym @ p
        alpha = rz / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        nr = float(r @ r) / nr0
        if nr <= tol * tol:
            return x, it, math.sqrt(nr)
        z = precond(r)
        rz_old = rz
        rz = float(r @ z)
        p = z + (rz / rz_old) * p


def ilu0(A):
    """[L, U] = ilu(A, struct('type', 'nofill')) as operator_divergence.m:
    243-247 calls it: incomplete LU without fill and without pivoting, row by
    row (the IKJ form: for every row i, eliminate with the earlier rows k of
    its own pattern, updating only entries in the pattern).  L is unit lower
    triangular, U upper triangular, (L U)_ij = A_ij on the pattern of A."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    N = A.shape[0]
    rows = [dict(zip(A.indices[A.indptr[i]:A.indptr[i + 1]], A.data[A.indptr[i]:A.indptr[i + 1]]))
            for i in range(N)]
    for i in range(N):
        ri = rows[i]
        for k in sorted(c for c in ri if c < i):
            ri[k] = ri[k] / rows[k][k]
            lik = ri[k]
            for j, ukj in rows[k].items():
                if j > k and j in ri:
                    ri[j] -= lik * ukj
    Li, Lj, Lv, Ui, Uj, Uv = [], [], [], [], [], []
    for i, ri in enumerate(rows):
        for j, v in ri.items():
            if j < i:
                Li.append(i); Lj.append(j); Lv.append(v)
            else:
                Ui.append(i); Uj.append(j); Uv.append(v)
    L = sp.csr_matrix((Lv, (Li, Lj)), shape=(N, N)) + sp.identity(N, format='csr')
    U = sp.csr_matrix((Uv, (Ui, Uj)), shape=(N, N))
    return L.tocsr(), U.tocsr()


def pcg_ilu(A, f, tol, maxit, L):
    """pressure_poisson.m:52-64: [dp, flag, norm1, iter] = pcg(-A, -f, CG_acc,
    CG_maxit, L, L', dp) with dp = 0: CG on -A with the preconditioner
    M = L L^T (L from ilu0(A)), stopping when the residual norm falls to
    tol ||f||.  Returns x, iterations."""
    K = (-sp.csr_matrix(A)).tocsr()
    b = -f
    Lt = L.T.tocsr()
    n2b = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    if n2b == 0.0:
        return x, 0
    r = b.copy()
    rho_old, p = None, None
    for it in range(1, maxit + 1):
        z = spla.spsolve_triangular(Lt, spla.spsolve_triangular(L, r, lower=True), lower=False)
        rho = float(r @ z)
        p = z.copy() if p is None else z + (rho / rho_old) * p
        q = K @ p
        alpha = rho / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        rho_old = rho
        if np.linalg.norm(r) <= tol * n2b:
            return x, it
    raise RuntimeError(f'pcg did not converge in {maxit} iterations')


def pressure_poisson(f, t, opts):
    """solvers/pressure/pressure_poisson.m:1-106."""
    g = opts.grid; s = opts.solversettings
    dp = np.zeros(g.Np)
    if s.poisson == 6:
        fhat = sfft.fft2(f.reshape(g.Npx, g.Npy, order='F'), workers=FFT_WORKERS)
        if np.max(np.abs(fhat)) < 1e-14:
            return dp
        phat = -fhat / s.Ahat
        return np.real(sfft.ifft2(phat, workers=FFT_WORKERS)).ravel(order='F')
    if np.max(np.abs(f)) < 1e-14:
        return dp
    if s.poisson == 1:
        return s.decomp(f)
    if s.poisson in (3, 4):
        x, it, nrm = cg_matlab(opts.discretization.A, f, s.CG_acc, dp, s.CG_maxit, s.A_pc)
        s.last_iters = it
        return x
    if s.poisson == 2:
        x, it = pcg_ilu(opts.discretization.A, f, s.CG_acc, s.CG_maxit, s.L)
        s.last_iters = it
        return x
    raise ValueError('wrong poisson method')


def pressure_additional_solve(V, p, t, opts):
    """solvers/pressure/pressure_additional_solve.m:1-26 (ydM at t for
    unsteady BC, :8-14)."""
    d = opts.discretization; g = opts.grid
    if opts.BC.BC_unsteady:
        set_bc_vectors(t, opts)
    _, R = F(V, V, p, t, opts)
    f = d.M @ (g.Om_inv * R) + d.ydM
    return p + pressure_poisson(f, t, opts)


def check_conservation(V, t, opts):
    """spatial/check_conservation.m:1-76 (boundary vectors at t for unsteady
    BC, :12-14)."""
    g = opts.grid; d = opts.discretization; BC = opts.BC; case = opts.case_obj
    if BC.BC_unsteady:
        set_bc_vectors(t, opts)
    uh, vh = V[:g.Nu], V[g.Nu:]
    uLe_i = case.uBC(g.x[0:1], g.yp, t, opts); uRi_i = case.uBC(g.x[-1:], g.yp, t, opts)
    vLo_i = case.vBC(g.xp, g.y[0:1], t, opts); vUp_i = case.vBC(g.xp, g.y[-1:], t, opts)
    maxdiv = np.max(np.abs(d.M @ V + d.yM))
    umom = np.sum(g.Omu * uh); vmom = np.sum(g.Omv * vh)
    if BC.u.left == 'dir': umom += np.sum(uLe_i * g.hy) * g.gx[0]
    if BC.u.right == 'dir': umom += np.sum(uRi_i * g.hy) * g.gx[-1]
    if BC.v.low == 'dir': vmom += np.sum(vLo_i * g.hx) * g.gy[0]
    if BC.v.up == 'dir': vmom += np.sum(vUp_i * g.hx) * g.gy[-1]
    kin = 0.5 * np.sum(g.Omu * uh ** 2) + 0.5 * np.sum(g.Omv * vh ** 2)
    if BC.u.left == 'dir': kin += 0.5 * np.sum(uLe_i ** 2 * g.hy) * g.gx[0]
    if BC.u.right == 'dir': kin += 0.5 * np.sum(uRi_i ** 2 * g.hy) * g.gx[-1]
    if BC.v.low == 'dir': kin += 0.5 * np.sum(vLo_i ** 2 * g.hx) * g.gy[0]
    if BC.v.up == 'dir': kin += 0.5 * np.sum(vUp_i ** 2 * g.hx) * g.gy[-1]
    return maxdiv, umom, vmom, kin


# ----------------------------------------------------------------------------
# time integration (unsteady/getRKmethod.m, time_ERK.m, solver_unsteady.m)
# ----------------------------------------------------------------------------


def get_rk_method(name):
    """Explicit tableaux of unsteady/getRKmethod.m (A, b, c)."""
    f = lambda *v: np.array(v, float)
    if name == 'FE11':
        A = f([0]); b = f(1)
    elif name == 'SSP22':
        A = f([0, 0], [1, 0]); b = f(.5, .5)
    elif name == 'SSP42':
        A = f([0, 0, 0, 0], [1 / 3, 0, 0, 0], [1 / 3, 1 / 3, 0, 0], [1 / 3, 1 / 3, 1 / 3, 0]); b = np.full(4, .25)
    elif name == 'SSP33':
        A = f([0, 0, 0], [1, 0, 0], [.25, .25, 0]); b = f(1 / 6, 1 / 6, 2 / 3)
    elif name == 'SSP43':
        A = f([0, 0, 0, 0], [.5, 0, 0, 0], [.5, .5, 0, 0], [1 / 6, 1 / 6, 1 / 6, 0]); b = f(1 / 6, 1 / 6, 1 / 6, .5)
    elif name == 'Wray3':
        A = np.zeros((3, 3)); A[1, 0] = 8 / 15; A[2, 0] = 8 / 15 - 17 / 60; A[2, 1] = 5 / 12
        b = f(8 / 15 - 17 / 60, 0, .75)
        c = f(0, A[1, 0], A[2, 0] + A[2, 1])
        return A, b, c
    elif name == 'RK33C2':
        A = f([0, 0, 0], [2 / 3, 0, 0], [1 / 3, 1 / 3, 0]); b = f(.25, 0, .75); return A, b, f(0, 2 / 3, 2 / 3)
    elif name == 'RK33P2':
        A = f([0, 0, 0], [1 / 3, 0, 0], [-1, 2, 0]); b = f(0, .75, .25); return A, b, f(0, 1 / 3, 1)
    elif name == 'RK44':
        A = f([0, 0, 0, 0], [.5, 0, 0, 0], [0, .5, 0, 0], [0, 0, 1, 0]); b = f(1 / 6, 1 / 3, 1 / 3, 1 / 6)
    elif name == 'RK44C2':
        A = f([0, 0, 0, 0], [.25, 0, 0, 0], [0, .5, 0, 0], [1, -2, 2, 0]); b = f(1 / 6, 0, 2 / 3, 1 / 6); return A, b, f(0, .25, .5, 1)
    elif name == 'RK44C23':
        A = f([0, 0, 0, 0], [.5, 0, 0, 0], [.25, .25, 0, 0], [0, -1, 2, 0]); b = f(1 / 6, 0, 2 / 3, 1 / 6); return A, b, f(0, .5, .5, 1)
    elif name == 'RK44P2':
        A = f([0, 0, 0, 0], [1, 0, 0, 0], [3 / 8, 1 / 8, 0, 0], [-1 / 8, -3 / 8, 1.5, 0]); b = f(1 / 6, -1 / 18, 2 / 3, 2 / 9); return A, b, f(0, 1, .5, 1)
    elif name == 'Heun33':
        A = f([0, 0, 0], [1 / 3, 0, 0], [0, 2 / 3, 0]); b = f(.25, 0, .75)
    elif name == 'Mid22':
        A = f([0, 0], [.5, 0]); b = f(0, 1); return A, b, f(0, .5)
    elif name == 'MTE22':
        A = f([0, 0], [2 / 3, 0]); b = f(.25, .75); return A, b, f(0, 2 / 3)
    elif name == 'RK56':
        A = f([0, 0, 0, 0, 0, 0], [.25, 0, 0, 0, 0, 0], [1 / 8, 1 / 8, 0, 0, 0, 0], [0, 0, .5, 0, 0, 0],
              [3 / 16, -3 / 8, 3 / 8, 9 / 16, 0, 0], [-3 / 7, 8 / 7, 6 / 7, -12 / 7, 8 / 7, 0])
        b = f(7 / 90, 0, 16 / 45, 2 / 15, 16 / 45, 7 / 90); return A, b, f(0, .25, .25, .5, .75, 1)
    elif name == 'DOPRI6':
        A = f([0, 0, 0, 0, 0, 0], [1 / 5, 0, 0, 0, 0, 0], [3 / 40, 9 / 40, 0, 0, 0, 0], [44 / 45, -56 / 15, 32 / 9, 0, 0, 0],
              [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729, 0, 0],
              [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656, 0])
        b = f(35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84)
    else:
        raise ValueError('RK method not available')
    A = np.atleast_2d(A)
    return A, b, A.sum(axis=1)


def time_ERK(Vn, pn, tn, dt, opts):
    """unsteady/time_ERK.m:1-173.  Unsteady BC: boundary vectors at tn (:62-64),
    F at each stage time (F.m:27-29), yM at the stage's end time for the
    divergence (:106-109), a pressure solve at every stage (:121-127), and at
    the end the additional solve at tn + dt or the last stage pressure
    (:145-164)."""
    g = opts.grid; d = opts.discretization
    unsteady = opts.BC.BC_unsteady
    A_RK, b_RK, c_RK = get_rk_method(opts.time.RK)
    s = len(b_RK)
    A_RK = np.vstack([A_RK[1:, :], b_RK[None, :]])
    c_RK = np.r_[c_RK[1:], 1.0]
    V = Vn.copy(); p = pn.copy()
    kV = np.zeros((g.NV, s))
    if unsteady:
        set_bc_vectors(tn, opts)
    yM = d.yM
    dp = pn
    for i in range(s):
        ti = tn + (0 if i == 0 else c_RK[i - 1] * dt)
        _, Frhs = F(V, V, p, ti, opts)
        kV[:, i] = g.Om_inv * (Frhs + d.G @ p)
        Vtemp = kV @ A_RK[i, :]
        ti = tn + c_RK[i] * dt
        if unsteady:
            set_bc_vectors(ti, opts)
            yM = d.yM
        f = (d.M @ (Vn / dt + Vtemp) + yM / dt) / c_RK[i]
        if unsteady or i > 0:
            dp = pressure_poisson(f, ti, opts)
        else:
            dp = pn
        V = Vn + dt * (Vtemp - c_RK[i] * g.Om_inv * (d.G @ dp))
    if unsteady and not opts.solversettings.p_add_solve:
        p = dp   # kp(:, end)
    else:
        p = pressure_additional_solve(V, p, tn + dt, opts)
    return V, p


# ----------------------------------------------------------------------------
# driver (main.m, input_check.m, solver_unsteady.m, process_iteration.m)
# ----------------------------------------------------------------------------


def _ns(**kw):
    return SimpleNamespace(**kw)


def build(case, overrides=None):
    """main.m:68-186: parameters -> options, mesh, BC, operators, IC, BC vectors."""
    prm = dict(case.parameters())
    if overrides:
        prm.update(overrides)
    opts = _ns()
    opts.case_obj = case
    opts.prm = prm
    opts.fluid = _ns(Re=float(prm['Re']))
    opts.grid = _ns(x1=prm.get('x1', 0.0), x2=prm.get('x2', 1.0), y1=prm.get('y1', 0.0), y2=prm.get('y2', 1.0),
                    Nx=prm['Nx'], Ny=prm['Ny'], sx=prm.get('sx', 1.0), sy=prm.get('sy', 1.0))
    opts.discretization = _ns(order4=int(prm.get('order4', 0)), alfa=float(prm.get('alfa', 81.0)), beta=float(prm.get('beta', 9 / 8)))
    opts.solversettings = _ns(poisson=int(prm.get('poisson', 1)), p_initial=int(prm.get('p_initial', 1)),
                              p_add_solve=int(prm.get('p_add_solve', 1)), CG_acc=float(prm.get('CG_acc', 1e-8)),
                              CG_maxit=int(prm.get('CG_maxit', 1000)))
    opts.time = _ns(RK=prm.get('RK', 'RK44'), dt=prm.get('dt'), t_start=float(prm.get('t_start', 0.0)),
                    t_end=prm.get('t_end'))
    x, y = case.mesh(opts)
    mesh_generation(opts, x, y)
    bct = case.BCtype()
    opts.BC = _ns(u=_ns(**bct['u']), v=_ns(**bct['v']), BC_unsteady=int(bct.get('BC_unsteady', 0)), p_inf=0.0)
    operators(opts)
    t = opts.time.t_start
    u0, v0, p0 = case.IC(t, opts)
    opts.V_start = np.r_[np.asarray(u0, float).ravel(order='F'), np.asarray(v0, float).ravel(order='F')]
    opts.p_start = np.asarray(p0, float).ravel(order='F')
    if opts.time.dt is None:
        # fixed dt from a CFL number on the initial field (case-defined)
        opts.time.dt = case.module.cfl_dt(opts, u0, v0)
    opts.time.dt = float(opts.time.dt)
    if opts.time.t_end is None:
        opts.time.t_end = opts.time.t_start + int(prm['nsteps']) * opts.time.dt
    opts.time.t_end = float(opts.time.t_end)
    set_bc_vectors(t, opts)
    return opts


def input_check(opts):
    """preprocessing/input_check.m:92-171: initial projection, initial p, maxres."""
    t = opts.time.t_start
    V = opts.V_start.copy()
    row = list(check_conservation(V, t, opts))
    if row[0] > 1e-12:
        d = opts.discretization
        f = d.M @ V + d.yM
        dp = pressure_poisson(f, t, opts)
        V = V - opts.grid.Om_inv * (d.G @ dp)
        row = list(check_conservation(V, t, opts))
    if opts.solversettings.p_initial == 1:
        p = pressure_additional_solve(V, opts.p_start.copy(), t, opts)
    else:
        p = opts.p_start.copy()
    maxres, _ = F(V, V, p, t, opts)
    return V, p, (1, opts.time.dt, t, maxres, *row)


def run(case, nsteps=None, overrides=None, callback=None):
    """solver_unsteady.m:72-133 + process_iteration.m:4,22,112-119.

    Returns (V, p, rows) where rows are the convergence.txt records
    (n, dt, t, res, maxdiv, umom, vmom, k)."""
    opts = build(case, overrides)
    V, p, row0 = input_check(opts)
    rows = [row0]
    dt = opts.time.dt
    nt = int(math.ceil((opts.time.t_end - opts.time.t_start) / dt - 1e-12)) if nsteps is None else nsteps
    t = opts.time.t_start
    n = 1
    for _ in range(nt):
        n += 1
        V, p = time_ERK(V, p, t, dt, opts)
        t = t + dt
        maxdiv, umom, vmom, kin = check_conservation(V, t, opts)
        maxres, _ = F(V, V, p, t, opts)
        rows.append((n, dt, t, maxres, maxdiv, umom, vmom, kin))
        if callback is not None:
            callback(n, V, p, t)
    return V, p, rows, opts

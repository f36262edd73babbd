"""Decaying 2-D turbulence on the periodic box [0,2pi]^2.

Random divergence-free initial field built from a streamfunction on the cell
corners, V = W^T psi, which is exactly discretely divergence-free
(postprocessing/get_velocity_from_streamfunction.m:1-29; used the same way by
case_files/decaying_turbulence/decaying_turbulence_IC.m:33).  psi has the
energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 10, with seeded uniform
phases; V is scaled so the domain-averaged kinetic energy 1/2 <|u|^2> is
``energy`` (default 1).

The pressure solve uses the exact Fourier symbol of the assembled
fourth-order A (``fft_symbol = 1``): the reference's FFT branch always takes
the second-order symbol (spatial/operators/operator_divergence.m:283), which
leaves the fourth-order projection with a divergence that grows until the run
diverges; ``fft_symbol = 0`` reproduces the reference's choice.
"""
import math

import numpy as np

from ._common import periodic, uniform_mesh, zeros_bc


def parameters():
    return dict(Re=1e4, Nx=4096, Ny=4096, x1=0.0, x2=2 * math.pi, y1=0.0, y2=2 * math.pi, sx=1.0, sy=1.0,
                order4=1, dt=None, CFL=0.4, t_start=0.0, nsteps=500, RK='RK44',
                poisson=6, fft_symbol=1, p_initial=1, p_add_solve=1, seed=2024, k0=10.0, energy=1.0)


def BCtype():
    return periodic()


mesh = uniform_mesh


def streamfunction(options):
    g = options.grid
    prm = options.prm
    Nx, Ny = g.Nx, g.Ny
    rng = np.random.default_rng(int(prm.get('seed', 2024)))
    kx = np.fft.fftfreq(Nx, d=1.0 / Nx)
    ky = np.fft.rfftfreq(Ny, d=1.0 / Ny)
    KX, KY = np.meshgrid(kx, ky, indexing='ij')
    k = np.sqrt(KX ** 2 + KY ** 2)
    k0 = float(prm.get('k0', 10.0))
    amp = np.zeros_like(k)
    nz = k > 0
    # |psi_hat|^2 ~ E(k) / k^3 with E(k) ~ k^4 exp(-2 (k/k0)^2)
    amp[nz] = np.sqrt(k[nz] * np.exp(-2.0 * (k[nz] / k0) ** 2))
    phase = np.exp(2j * np.pi * rng.random(k.shape))
    psi = np.fft.irfft2(amp * phase, s=(Nx, Ny))
    return psi


def velocity_from_streamfunction(options, psi):
    """V = W^T psi with W the discrete curl on the periodic staggered grid
    (get_velocity_from_streamfunction.m:9-29)."""
    g = options.grid
    hx = np.asarray(g.hx); hy = np.asarray(g.hy)
    # u = d psi/dy  (W1D: rows j of -1/hy(j-1) at j-1 ... periodic wrap)
    # W1D = spdiags([-1/h 1/h],[-1 0]); W1D(1,end) = -W1D(1,1); V = W^T psi
    # (W^T psi)_u(i,j) = -( psi(i,j)/hy(j) - psi(i,j+1)/hy(j) ) with wrap
    psi = psi.reshape(g.Nx, g.Ny)
    u = -(psi / hy[None, :] - np.roll(psi, -1, axis=1) / hy[None, :])
    v = (psi / hx[:, None] - np.roll(psi, -1, axis=0) / hx[:, None])
    return u, v


def IC(t, options):
    g = options.grid
    psi = streamfunction(options)
    u, v = velocity_from_streamfunction(options, psi)
    e = 0.5 * (np.mean(u ** 2) + np.mean(v ** 2))
    s = math.sqrt(float(options.prm.get('energy', 1.0)) / e)
    return u * s, v * s, np.zeros((g.Npx, g.Npy))


def cfl_dt(options, u, v):
    h = min(np.min(options.grid.hx), np.min(options.grid.hy))
    umax = max(np.max(np.abs(u)), np.max(np.abs(v)))
    return float(options.prm.get('CFL', 0.4)) * h / umax


uBC = zeros_bc
vBC = zeros_bc

"""Throughput of the INS2D explicit-RK time stepping on one MI355X.

Workload (default): decaying 2-D turbulence on the periodic box [0,2pi]^2,
4096 x 4096 cells, fourth-order energy-conserving discretization (alfa = 81,
beta = 9/8), a seeded divergence-free random initial field with energy
spectrum E(k) ~ k^4 exp(-2 (k/10)^2) scaled to unit kinetic energy, Re = 1e4,
classical RK4 with a fixed time step from CFL 0.4, FFT pressure solve with the
exact symbol of the fourth-order A, FP64.  ``--case taylor_green`` runs the
second-order Taylor-Green vortex at 1024 x 1024 (Re = 2000, CFL 0.5).
A "step" is one RK4 time step (4 stages) of the whole field plus the
additional pressure solve; the timed region excludes setup and diagnostics.

metric: Mcell-RK-stages/s = cells x RK stages x steps / seconds / 1e6.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the path runs on one
GPU (a sharded layout, row slabs of the grid with halo exchange and a
distributed transpose for the FFT solve, is designed but not built).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def progress(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--case', default='decaying_turbulence')
    ap.add_argument('--N', type=int, default=0)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-steps', type=int, default=10)
    ap.add_argument('--members', type=int, default=512)
    ap.add_argument('--train-steps', type=int, default=4000)
    ap.add_argument('--pairs', type=int, default=20000)
    return ap.parse_args()


def cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(case_args, steps, timeout_s=240):
    """The CPU oracle (numpy/scipy restatement of the reference algorithm)
    timed on a bounded sample of the same workload, in a child process with
    its own time limit (BLAS single-threaded; the FFT solve on the host's
    cores)."""
    import subprocess
    env = dict(os.environ)
    for k in ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS', 'MKL_NUM_THREADS'):
        env[k] = '1'
    cmd = [sys.executable, '-u', os.path.abspath(__file__), '--cpu-child', json.dumps(case_args), str(steps)]
    try:
        # the child's progress lines go straight to this process's stderr
        out = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=timeout_s)
        return json.loads(out.stdout.strip().splitlines()[-1])
    except subprocess.TimeoutExpired:
        return {'value': None, 'unit': 'Mcell-RK-stages/s', 'cores': 1, 'kind': 'port',
                'sample': f'CPU oracle sample did not finish within {timeout_s} s'}


def cpu_child(case_json, steps):
    progress('cpu oracle child started')
    from ins2d_amd.cases import load
    from oracle import ins2d_oracle as O
    case_args = json.loads(case_json)
    case = load(**case_args)
    # the FFT pressure solve on all host cores (at most 16), the sparse
    # operator products single-threaded (scipy)
    O.FFT_WORKERS = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    opts = O.build(case)
    V, p, _ = O.input_check(opts)
    t_setup = time.perf_counter() - t0
    progress(f'cpu oracle setup {t_setup:.1f} s')
    dt = opts.time.dt
    # at least `steps` steps and, for fast cases, on until about 10 s of CPU
    # work (a sample long enough to time)
    t0 = time.perf_counter()
    t_print = t0
    n = 0
    while n < steps or (time.perf_counter() - t0 < 10.0 and n < 200 * steps):
        V, p = O.time_ERK(V, p, n * dt, dt, opts)
        n += 1
        if time.perf_counter() - t_print >= 1.0 or n == steps:
            progress(f'cpu oracle step {n}')
            t_print = time.perf_counter()
    steps = n
    el = time.perf_counter() - t0
    stages = len(O.get_rk_method(opts.time.RK)[1])
    cells = opts.grid.Np
    solver = {1: 'direct sparse LU', 6: 'FFT', 3: 'IC(0)-PCG', 4: 'IC(0)-PCG'}[opts.solversettings.poisson]
    fft = opts.solversettings.poisson == 6
    cores = O.FFT_WORKERS if fft else 1
    threads = (f'scipy.fft on {cores} threads, sparse products single-threaded' if fft
               else 'numpy/scipy single thread')
    print(json.dumps({'value': cells * stages * steps / el / 1e6, 'unit': 'Mcell-RK-stages/s', 'cores': cores,
                      'kind': 'port',
                      'sample': f'{steps} RK steps of the {opts.grid.Nx}x{opts.grid.Ny} case, {solver} pressure '
                                f'({el:.1f} s, setup {t_setup:.1f} s excluded), {threads}, '
                                f'CPU {cpu_model()}, nproc {os.cpu_count()}'}))


def copy_calibration():
    """A plain 256 MiB device-to-device copy, three times: the reference
    transfer for scaling the FETCH_SIZE / WRITE_SIZE hardware counters."""
    import torch
    n = 64 * 1024 * 1024
    x = torch.ones(n, dtype=torch.float32, device='cuda')
    torch.cuda.synchronize()
    for _ in range(3):
        y = x.clone()
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes': 4 * n, 'copies': 3, 'check': float(y[-1])}))


def measured_traffic(key):
    """HBM bytes per launch of the roofline kernel from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (profiles/pmc_traffic.json)."""
    try:
        with open(os.path.join(ROOT, 'profiles', 'pmc_traffic.json')) as fh:
            return json.load(fh).get(key, {}).get('bytes_per_launch')
    except (OSError, ValueError):
        return None


def bench_ensemble(a):
    """Batched ensemble: 512 cavity/channel members of uneven size (48..512
    cells on the long side, aspect 1..4), Re in [100, 5000], speed in
    [0.5, 1.5], RK4, PCG pressure to 1e-10 with the fast-diagonalization
    inverse of each member's pressure matrix as preconditioner (the cases
    ensemble_2level and ensemble_jacobi use the two-level and the plain
    Jacobi preconditioner)."""
    from ins2d_amd import host
    from ins2d_amd.cases import ensemble
    precond = {'ensemble_jacobi': 'jacobi', 'ensemble_2level': 'two_level'}.get(a.case, 'fdm')
    members = ensemble.members(a.members, precond=precond)
    t0 = time.perf_counter()
    B = host.Batch(members)
    B.input_check(1)
    tile = B.autotune()
    progress('stage tile shape', tile)
    t_setup = time.perf_counter() - t0
    B.step(a.warmup)
    it0, _, cells = B.stats()
    t0 = time.perf_counter()
    B.step(a.steps)
    el = time.perf_counter() - t0
    it1, itmax, _ = B.stats()
    value = cells * 4 * a.steps / el / 1e6
    out = {'metric': 'Mcell-RK-stages/s per GPU', 'value': value, 'unit': 'Mcell-RK-stages/s', 'n_gpus': 1,
           'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': 1e3 * el / a.steps, 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
           'config': {'workload': f'ensemble of {a.members} cavity/channel members, RK44, '
                                  + {'jacobi': 'Jacobi', 'two_level': 'two-level (Jacobi + coarse grid)',
                                     'fdm': 'fast-diagonalization'}[precond] + '-PCG 1e-10',
                      'cells': cells, 'setup_s': t_setup, 'stage_tile': list(tile),
                      'pcg_iterations_per_member_solve': (it1 - it0) / (a.members * 4 * a.steps)}}
    B.close()
    print(json.dumps(out))


def bench_poisson_learned(a):
    """Per-solve pressure time, learned vs plain, on the backward-facing step
    1024x256 (tanh-stretched, Re 800 grid), PCG to a true relative residual
    of 1e-10.  Training pairs (a.pairs, default 20,000) come from plain-solver
    runs at Re 200..1200, generated into device memory with the coarse-grid
    solution of each right-hand side; the held-out right-hand sides from runs
    at Re 300, 700, 1100 at later times.  Variants: plain Jacobi-PCG from
    x0 = 0; the two-level PCG from x0 = 0; the two-level PCG started from the
    coarse-grid solution (the same method without the network); learned =
    the network's correction of the coarse-grid solution as the start of the
    two-level PCG."""
    from ins2d_amd import host, learn
    Nx, Ny = (a.N or 1024), (a.N // 4 if a.N else 256)
    Re_train = [200.0, 400.0, 600.0, 800.0, 1000.0, 1200.0]
    per_run = max(1, a.pairs // len(Re_train))
    n = per_run * len(Re_train)
    batch = 4
    t0 = time.perf_counter()
    net = learn.make_net(Nx, Ny, out_scale=1.0, seed=0)
    net.data_reserve(n + batch // 2)
    learn.generate_pairs_dev(net, Nx, Ny, Re_train, per_run, every=3, skip=100, progress=progress)
    out_scale = learn.out_scale_from_net(net, n)
    net.init(0, out_scale)
    held = learn.generate_pairs(Nx, Ny, [300.0, 700.0, 1100.0], steps=340, every=10, skip=400, progress=progress)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = learn.train_stored(net, n, steps=a.train_steps, batch=batch, lr=2e-3, progress=progress, every=250)
    t_train = time.perf_counter() - t0
    solver, _, _, _ = host.setup('BFS_unsteady', Nx=Nx, Ny=Ny, poisson=3, CG_acc=1e-10, CG_maxit=100000)
    rhs = [h[0] for h in held]
    # accuracy of the start vectors on the held-out set (relative L2 error)
    solver.set_poisson_accel(None, two_level=True)
    err_c, err_n = [], []
    for f, p, *_ in held[::10]:
        xc = solver.poisson_coarse(f)
        pn = net.forward(f.reshape(Ny, Nx), xc.reshape(Ny, Nx)).ravel()
        err_c.append(np.linalg.norm(xc - p) / np.linalg.norm(p))
        err_n.append(np.linalg.norm(pn - p) / np.linalg.norm(p))
    progress(f'held-out start error: coarse {np.median(err_c):.3e}, learned {np.median(err_n):.3e}')
    variants = [('plain', {}), ('two_level', dict(two_level=True)),
                ('coarse_start', dict(two_level=True, coarse_start=True)),
                ('learned', dict(warm_start=True, two_level=True, coarse_input=True))]
    res = learn.evaluate(solver, net, rhs, variants, progress=progress)
    med = {k: float(np.median(v['ms'])) for k, v in res.items()}
    out = {'metric': 'Poisson time learned vs plain PCG @1e-10 residual', 'value': med['plain'] / med['learned'],
           'unit': 'x (median plain / median learned per-solve GPU time)', 'n_gpus': 1, 'steps': len(rhs),
           'warmup': 1, 'ms_per_step': med['learned'], 'higher_is_better': True, 'scaling': 'weak',
           'vs_baseline': None, 'dtype': 'f64 (solver), bf16/fp32 (network)', 'data': 'synthetic',
           'config': {'workload': f'backward-facing step {Nx}x{Ny} tanh-stretched, pressure RHS from plain runs',
                      'training_pairs': n, 'held_out_rhs': len(rhs), 'train_steps': a.train_steps,
                      'train_batch': batch, 'train_s': t_train, 'data_s': t_data, 'out_scale': out_scale,
                      'loss_first': float(np.mean(losses[:10])), 'loss_last': float(np.mean(losses[-50:])),
                      'held_out_rel_error_coarse': float(np.median(err_c)),
                      'held_out_rel_error_learned': float(np.median(err_n))},
           'variants': {k: {'median_ms': med[k], 'median_iterations': float(np.median(v['it'])),
                            'max_true_relres': float(np.max(v['rel'])), 'fallbacks': v['fallbacks']}
                        for k, v in res.items()},
           'network_gain_over_two_level': med['two_level'] / med['learned'],
           'network_gain_over_coarse_start': med['coarse_start'] / med['learned']}
    agg = solver.accel_stats()[1]
    out['config']['aggregate'] = agg
    solver.close()
    net.close()
    print(json.dumps(out))


def stage_fields(RK):
    """FP64 fields per cell that each stage's momentum-RHS launch must move:
    read V (2), write Vtemp (2), write k_i when a later stage reads it (2),
    read the k_j of the shifted tableau row (2 each; time_ERK.m:29,95-101)."""
    from ins2d_amd import host
    A, b, c = host.get_rk_method(RK)
    s = len(b)
    At = np.vstack([A[1:], b[None, :]])
    keep = [any(At[i, j] != 0 for i in range(j + 1, s)) for j in range(s)]
    return [2 + 2 + (2 if keep[i] else 0) + 2 * sum(1 for j in range(i) if At[i, j] != 0) for i in range(s)]


def kernel_table(raw, RK, poisson, cells, Nx, assembled=False):
    """Per kernel class: launches, time per launch and achieved bandwidth on
    algorithmic bytes (8 B per FP64 value a launch must read or write once
    per cell), weighted over the launches the profile pass made.  raw maps
    the library's profile kinds to (ms, launches): 0-3 the RK stages'
    momentum RHS, divergence, Poisson solve and projection, 4-6 the RHS,
    divergence and solve of the additional solve.  assembled: the
    fourth-order operators with walls in CSR form, whose HBM traffic is set by
    the stored operator entries, not by the fields (no byte model)."""
    fields = stage_fields(RK)
    s = len(fields)
    fft = poisson == 6
    # momentum RHS: the stage launches at their tableau rows, the additional
    # solve's F (read V, p, write F: 5 fields)
    n0, n4 = raw[0][1], raw[4][1]
    steps = n0 // s if s else 0
    rhs_fields = (steps * sum(fields) + n4 * 5) / max(n0 + n4, 1)
    # divergence: stage f = M (Vn/dt + Vtemp) (5 fields), additional f = M F (3)
    n1, n5 = raw[1][1], raw[5][1]
    div_fields = (n1 * 5 + n5 * 3) / max(n1 + n5, 1)
    # FFT solve: lower bound read f, write dp (2 fields); with the divergence
    # fused the stage solve reads Vn, Vtemp in place of f (5) and the
    # additional solve reads F and p and writes p (4); else the additional
    # solve reads f and p and writes p (3)
    n2, n6 = raw[2][1], raw[6][1]
    # the divergence computed inside the forward row transform (second-order
    # periodic kernels; fourth order with 4096-point rows): no divergence
    # launches of their own
    fused = fft and raw[8][1] > 0 and raw[1][1] + raw[5][1] == 0
    if fused:
        sol_fields = (n2 * 5 + n6 * 4) / max(n2 + n6, 1)
        sol_model = 'stage: read Vn, Vtemp, write dp (divergence fused); additional: read F, p, write p'
    elif fft:
        sol_fields = (n2 * 2 + n6 * 3) / max(n2 + n6, 1)
        sol_model = 'stage: read f, write dp; additional: read f, p, write p (lower bound: 3 kernels)'
    else:
        sol_fields = None
        sol_model = 'iterative or matrix solve: not an HBM stream'
    table = {
        'momentum_rhs': ((raw[0][0] + raw[4][0], n0 + n4), rhs_fields,
                         'momentum RHS + stage combination (k_rhs_*)',
                         f'stages read V + tableau k_j, write Vtemp (+ k_i): {fields} fields; additional: read V, p, '
                         'write F: 5 fields; launch-weighted'),
        'divergence': ((raw[1][0] + raw[5][0], n1 + n5), div_fields, 'divergence RHS (k_div_*)',
                       'stage: read Vn, Vtemp, write f; additional: read F, write f; launch-weighted'),
        'poisson': ((raw[2][0] + raw[6][0], n2 + n6), sol_fields, 'pressure solve', sol_model),
        'projection': (raw[3], 7, 'projection (k_proj_*)', 'read Vn, Vtemp, dp, write V'),
    }
    single = {'momentum_rhs', 'divergence', 'projection'}
    if assembled:
        note = 'assembled CSR operators: the stored entries dominate the traffic (not modelled)'
        for k in ('momentum_rhs', 'divergence', 'projection'):
            times, _, kname, _ = table[k]
            table[k] = (times, None, kname + ' [CSR, k_w4_*]', note)
        single.discard('momentum_rhs')   # face products + right-hand side: two kernels
    if fft and raw[8][1]:
        # the kernels of the register-transform FFT solve; the half spectrum
        # holds Nx/2 + 1 complex values per row of Nx cells
        spec = 2.0 * (Nx // 2 + 1) / Nx
        if fused:
            fwd = (n2 * (4 + spec) + n6 * (2 + spec)) / max(n2 + n6, 1)
            fwd_model = 'stage: read Vn, Vtemp (divergence fused), write the half spectrum; additional: read F'
        else:
            fwd = 1 + spec
            fwd_model = 'read f, write the half spectrum'
        table['fft_rows_forward'] = (raw[8], fwd, 'FFT forward row transform (k_fftr_r2c_rows)', fwd_model)
        table['fft_cols_solve'] = (raw[9], 2 * spec, 'FFT column transforms + symbol (k_fftr_cols_solve)',
                                   'read and write the half spectrum')
        table['fft_rows_inverse'] = (raw[10], spec + 1, 'FFT inverse row transform (k_fftr_c2r_rows)',
                                     'read the half spectrum, write dp')
        table['fft_rows_inverse_add'] = (raw[11], spec + 2, 'FFT inverse row transform, p += dp (k_fftr_c2r_rows)',
                                         'read the half spectrum and p, write p')
        single |= {'fft_rows_forward', 'fft_cols_solve', 'fft_rows_inverse', 'fft_rows_inverse_add'}
    out = {}
    for name, ((ms, n), f, kname, model) in table.items():
        us = 1e3 * ms / n if n else None
        bpc = 8.0 * f if f is not None else None
        gbs = bpc * cells / (us * 1e-6) / 1e9 if (us and bpc) else None
        out[name] = {'kernel': kname, 'launches': n, 'ms_total': ms, 'us_per_launch': us, 'bytes_per_cell': bpc,
                     'achieved_GBs': gbs, 'frac': gbs / HBM_PEAK_GBS if gbs else None, 'byte_model': model,
                     'single_kernel': name in single}
    return out


CASES = {
    # workload name -> (case plugin, overrides, description)
    'taylor_green': ('taylor_green', lambda N: dict(Nx=N or 1024, Ny=N or 1024),
                     'Taylor-Green + noise, periodic {Nx}x{Ny}, order 2, RK44, FFT pressure, Re=2000'),
    'decaying_turbulence': ('decaying_turbulence', lambda N: dict(Nx=N or 4096, Ny=N or 4096, order4=1, fft_symbol=1),
                            'decaying turbulence, periodic {Nx}x{Ny}, order 4, RK44, FFT pressure (exact symbol), Re=1e4'),
    'LDC_unsteady': ('LDC_unsteady', lambda N: dict(Nx=N or 64, Ny=N or 64, Re=1000.0, dt=1e-3, RK='RK44', poisson=3,
                                                     CG_acc=1e-10, CG_maxit=20000),
                     'lid-driven cavity {Nx}x{Ny}, order 2, RK44, dt=1e-3, Jacobi-PCG 1e-10, Re=1000'),
    'LDC_unsteady_2level': ('LDC_unsteady', lambda N: dict(Nx=N or 64, Ny=N or 64, Re=1000.0, dt=1e-3, RK='RK44',
                                                            poisson=3, CG_acc=1e-10, CG_maxit=20000,
                                                            CG_precond='two_level'),
                            'lid-driven cavity {Nx}x{Ny}, order 2, RK44, dt=1e-3, two-level PCG 1e-10, Re=1000'),
    'LDC_unsteady_ic0': ('LDC_unsteady', lambda N: dict(Nx=N or 64, Ny=N or 64, Re=1000.0, dt=1e-3, RK='RK44', poisson=4,
                                                         CG_acc=1e-10, CG_maxit=20000, CG_precond='ic0'),
                         'lid-driven cavity {Nx}x{Ny}, order 2, RK44, dt=1e-3, IC(0)-PCG 1e-10 (the reference preconditioner), '
                         'Re=1000'),
    'LDC_unsteady_direct': ('LDC_unsteady', lambda N: dict(Nx=N or 64, Ny=N or 64, Re=1000.0, dt=1e-3, RK='RK44',
                                                            poisson=1),
                            'lid-driven cavity {Nx}x{Ny}, order 2, RK44, dt=1e-3, direct pressure solve '
                            '(fast diagonalization), Re=1000'),
    'LDC_unsteady_fdm': ('LDC_unsteady', lambda N: dict(Nx=N or 64, Ny=N or 64, Re=1000.0, dt=1e-3, RK='RK44', poisson=3,
                                                      CG_acc=1e-10, CG_maxit=20000, CG_precond='fdm'),
                         'lid-driven cavity {Nx}x{Ny}, order 2, RK44, dt=1e-3, PCG 1e-10 with the fast-diagonalization '
                         'preconditioner, Re=1000'),
    'LDC_inviscid': ('LDC_inviscid', lambda N: dict(Nx=N or 64, Ny=N or 64, dt=0.005),
                     'inviscid lid-driven cavity {Nx}x{Ny}, order 4 with walls (assembled operators), RK44, '
                     'direct pressure solve (dense inverse), Re=1e100'),
    'BFS_unsteady': ('BFS_unsteady', lambda N: dict(Nx=N or 1024, Ny=(N // 4 if N else 256)),
                     'backward-facing step {Nx}x{Ny} tanh-stretched, order 2, RK44, direct pressure solve '
                     '(fast diagonalization), Re=800'),
    'BFS_unsteady_pcg': ('BFS_unsteady', lambda N: dict(Nx=N or 1024, Ny=(N // 4 if N else 256), poisson=3),
                         'backward-facing step {Nx}x{Ny} tanh-stretched, order 2, RK44, Jacobi-PCG 1e-10, Re=800'),
    'BFS_unsteady_2level': ('BFS_unsteady', lambda N: dict(Nx=N or 1024, Ny=(N // 4 if N else 256), poisson=3,
                                                           CG_precond='two_level'),
                            'backward-facing step {Nx}x{Ny} tanh-stretched, order 2, RK44, two-level PCG 1e-10, Re=800'),
}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({'error': 'this path runs on one GPU; sharding over several GPUs is designed, not built',
                          'n_gpus': a.gpus}))
        sys.exit(2)
    if a.case in ('ensemble', 'ensemble_jacobi', 'ensemble_2level'):
        return bench_ensemble(a)
    if a.case == 'poisson_learned':
        return bench_poisson_learned(a)
    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:
        have_torch = False
    from ins2d_amd import host

    plugin, mk, desc = CASES[a.case]
    over = mk(a.N)
    progress('setup', a.case, over)
    solver, opts, V0, p0 = host.setup(plugin, **over)
    solver.set_state(V0, p0)
    solver.input_check(0.0, 1)
    progress('input_check done')
    dt = opts.dt
    RK = opts.prm.get('RK', 'RK44')
    stages = len(host.get_rk_method(RK)[1])
    cells = solver.Np
    t = 0.0
    solver.step(t, dt, a.warmup)
    t += a.warmup * dt
    solver.sync()
    if have_torch:
        torch.cuda.synchronize()
    it0 = solver.solver_stats()[2]
    progress('warmup done; timing', a.steps, 'steps')
    t0 = time.perf_counter()
    solver.step(t, dt, a.steps)
    solver.sync()
    if have_torch:
        torch.cuda.synchronize()
    el = time.perf_counter() - t0
    it1 = solver.solver_stats()[2]
    t += a.steps * dt
    value = cells * stages * a.steps / el / 1e6

    progress(f'timed region {el:.3f} s')
    # the same stepping with the reference's per-step diagnostics
    # (process_iteration.m:4,22: maxres of an extra F, maxdiv, momentum, energy)
    nd = max(5, a.steps // 4)
    t0 = time.perf_counter()
    for _ in range(nd):
        solver.step(t, dt, 1)
        t += dt
        solver.diagnostics(t)
    el_diag = time.perf_counter() - t0
    value_diag = cells * stages * nd / el_diag / 1e6
    progress(f'with per-step diagnostics: {1e3 * el_diag / nd:.3f} ms/step')
    # per-kernel timing with HIP events on the library's stream (separate pass)
    nprof = max(2, min(a.steps, 50))
    solver.profile(True)
    solver.step(t, dt, nprof)
    solver.sync()
    # launch classes of the library's profiler: the RK stages (0-3) and the
    # additional solve at the end of each step (4-6)
    kinds = {0: 'momentum_rhs', 1: 'divergence', 2: 'poisson', 3: 'projection',
             4: 'momentum_rhs_additional', 5: 'divergence_additional', 6: 'poisson_additional',
             8: 'fft_rows_forward', 9: 'fft_cols_solve', 10: 'fft_rows_inverse', 11: 'fft_rows_inverse_add'}
    raw = {}
    for k, nm in kinds.items():
        ms, n = solver.profile_read(k)
        raw[k] = (ms, n)
    solver.profile(False)
    t += nprof * dt
    diag = solver.diagnostics(t)
    assembled = bool(opts.prm.get('order4', 0)) and solver.path == 0
    kernels = kernel_table(raw, RK, opts.prm.get('poisson', 6), cells, opts.grid.Nx, assembled)
    # the dominant kernel: the most time per step among single kernels (the
    # pressure solve is a class of several kernels when it is an FFT)
    dom = max((v for v in kernels.values() if v['ms_total'] and v['single_kernel']), key=lambda v: v['ms_total'])
    achieved = dom['achieved_GBs']

    out = {
        'metric': 'Mcell-RK-stages/s per GPU',
        'value': value,
        'unit': 'Mcell-RK-stages/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': 1e3 * el / a.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': desc.format(Nx=opts.grid.Nx, Ny=opts.grid.Ny) + f', dt={dt:.4e}',
                   'cells': cells, 'rk_stages': stages, 'kernel_path': solver.path},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': (achieved / HBM_PEAK_GBS) if achieved else None,
                     'traffic': measured_traffic(f'{a.case}_{opts.grid.Nx}') if opts.grid.Nx == opts.grid.Ny else None,
                     'traffic_unit': 'HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)',
                     'algorithmic_bytes_per_launch': (dom['bytes_per_cell'] * cells
                                                      if dom['bytes_per_cell'] is not None else None),
                     'kernel': dom['kernel'],
                     'bytes_per_cell': dom['bytes_per_cell'],
                     'us_per_launch': dom['us_per_launch'],
                     'byte_model': dom['byte_model']},
        'with_diagnostics': {'value': value_diag, 'ms_per_step': 1e3 * el_diag / nd, 'steps': nd},
        'kernels': kernels,
        'diagnostics': {'maxres': diag[0], 'maxdiv': diag[1], 'k': diag[4]},
    }
    s = stages
    if it1 > it0:
        its = (it1 - it0) / (a.steps * (s - 1 + 1))
        out['config']['pcg_iterations_per_solve'] = its
        us_solve = kernels['poisson']['us_per_launch']
        if us_solve and opts.prm.get('CG_precond', 'jacobi') == 'jacobi' and cells > 32768:
            # the multi-block Jacobi-PCG iteration (two launches): read z and p,
            # write p and q; read p, q, x, r and 1/diag(A), write x, r, z
            us_it = us_solve / its
            gbs = 96.0 * cells / (us_it * 1e-6) / 1e9
            out['pcg_iteration'] = {'us_per_iteration': us_it, 'bytes_per_cell': 96, 'achieved_GBs': gbs,
                                    'frac': gbs / HBM_PEAK_GBS,
                                    'byte_model': 'k_cg2_apply: read z, p, write p, q; k_cg2_update: read p, q, '
                                                  'x, r, 1/diag(A), write x, r, z'}
    solver.close()
    progress('profile pass done')
    if not a.no_cpu_baseline:
        progress('cpu baseline')
        cpu_over = dict(over)
        if opts.prm.get('poisson', 6) != 6:
            # wall-bounded cases: the reference's own configuration solves the
            # pressure directly (LDC_unsteady/BFS_unsteady parameters files)
            cpu_over['poisson'] = 1
        if plugin == 'decaying_turbulence':
            # the CPU sample runs the same case on a 1024^2 grid (per-cell cost)
            cpu_over.update(Nx=min(opts.grid.Nx, 1024), Ny=min(opts.grid.Ny, 1024))
        out['cpu_baseline'] = cpu_baseline(dict(name=plugin, **cpu_over), a.cpu_steps)
    print(json.dumps(out))


if __name__ == '__main__':
    if len(sys.argv) > 1 and sys.argv[1] == '--cpu-child':
        cpu_child(sys.argv[2], int(sys.argv[3]))
    elif len(sys.argv) > 1 and sys.argv[1] == '--copy-calibration':
        copy_calibration()
    else:
        main()

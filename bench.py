"""Benchmark of the FFT-VIE hot path on one MI355X.

Workload (the headline metric's configuration): synthetic head-like phantom,
three nested ellipsoids (semi-axes 0.09/0.08/0.11 m, shells 1/0.85/0.7,
per-shell eps_r ~ U[20,80], sigma ~ U[0.1,2] S/m, seed 0) scaled by 0.85 so
the outer shell fits the 128^3 grid at 1.5 mm (0.192 m), 298 MHz, JVIE
form II, complex128, 256^3 x 6 circulant assembled on the device.

A step is one matvec y = A_N x of the JVIE operator over the compact
non-air DOFs (mv_AN 'notransp'), x and y resident in HBM. `value` is FP64
matvecs/s; the same line carries FP32 matvecs/s, time-to-1e-6 (true
residual) for plain FP64 GMRES(50) and for complex64 matvecs in
block-orthogonalised cycles with FP64 iterative refinement (GMRES mode 3),
the refined solution's error against the FP64 solution, the
per-kernel HIP-event timings with their HBM rates, the roofline of the
dominant kernel and a CPU baseline (the oracle's numpy/scipy restatement of
mv_AN on a bounded sample).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128]
                       [--no-solve] [--no-cpu] [--profile-only]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def kernel_bytes(L, M, N, Px, Py, Pz, nidx, s):
    """Compulsory HBM bytes of each matvec kernel (s = bytes per complex).

    fN is counted as stored: an assembled operator has the circulant's
    per-axis parity and is kept as its octant [6][Pz/2+1][Py/2+1][Px/2+1].
    """
    n = L * M * N
    A = 3 * Px * M * N * s
    B = 3 * Px * Py * N * s
    fN = 6 * (Px // 2 + 1) * (Py // 2 + 1) * (Pz // 2 + 1) * s
    pos = 3 * n * 4
    return {
        'fwd_x': pos + nidx * s + A,
        'fwd_y': A + B,
        'z_product': 2 * B + fN,
        'inv_y': B + A,
        'inv_x_epilogue': A + pos + 2 * nidx * s + 2 * n * s,
    }


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo)."""
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or 'unknown'


def solve_bytes(cycle_lengths, nidx, bmv, s):
    """Algorithmic HBM bytes of a restarted GMRES solve: per Arnoldi step k of
    a cycle, one product (bmv) plus classical Gram-Schmidt over k basis
    vectors, s (2k + 3) nidx (the multi-dot reads V(:,1:k) and w, the
    combination reads them again and writes w, the new basis vector is
    written)."""
    tot = 0
    for m in cycle_lengths:
        for k in range(1, m + 1):
            tot += bmv + s * (2 * k + 3) * nidx
    return tot


def block_solve_bytes(cycle_lengths, nidx, bmv, s, sblk=8):
    """Algorithmic HBM bytes of a refined solve with block-orthogonalised
    cycles (GMRES mode 3): per product bmv plus its local Gram-Schmidt
    against the j + 1 columns of its block, s (2 (j + 1) + 3) nidx (dots read
    the columns and w, the combination reads them and w again and writes the
    new column); per block of sb products with k basis columns before it,
    s (2 k + 6 sb) nidx (the block's dots read k + sb columns, its Gram sb,
    the projection k + sb and writes sb, the second CholeskyQR reads and
    writes sb)."""
    tot = 0
    for m in cycle_lengths:
        k = 1
        done = 0
        while done < m:
            sb = min(sblk, m - done)
            for j in range(sb):
                tot += bmv + s * (2 * (j + 1) + 3) * nidx
            tot += s * (2 * k + 6 * sb) * nidx
            k += sb
            done += sb
    return tot


def other_workloads(torch, dev, geometry, lib, vie, cpu=True, names=None):
    """The 32^3 sphere (radius 0.1 m, eps_r 50, sigma 0.5 S/m, 6.25 mm, 298 MHz)
    and the 96x96x112 head phantom at 2 mm (128 and 298 MHz): GPU FP64
    matvecs/s and GMRES(50) time to 1e-6, beside the CPU restatement's
    (oracle mv_AN + pgmres, numpy + scipy.fft) matvecs/s and, where its
    estimated duration stays under 200 s, its time to 1e-6 (cpu=False: the GPU
    figures only)."""
    from oracle import krylov as okr, mvp as omvp     # CPU baseline only
    cores = min(os.cpu_count() or 1, 16)
    rows = []
    for name, shape, dx, freq in (('sphere_32cubed_6p25mm_298MHz', (32, 32, 32), 6.25e-3, 298e6),
                                  ('head_96x96x112_2mm_128MHz', (96, 96, 112), 2e-3, 128e6),
                                  ('head_96x96x112_2mm_298MHz', (96, 96, 112), 2e-3, 298e6)):
        if names is not None and name not in names:
            continue
        L, M, N = shape
        em = vie.EMconstants(freq)
        if name.startswith('sphere'):
            r = geometry.cell_grid(L, M, N, dx)
            inside = np.linalg.norm(r, axis=-1) < 0.1
            e_r = np.where(inside, 50.0 - 1j * 0.5 / (em['eo'] * em['omega']), 1.0 + 0j)
        else:
            r, eps, sig, _ = geometry.head_phantom(L, M, N, dx, seed=0, scale=1.0)
            e_r = eps - 1j * sig / (em['eo'] * em['omega'])
        Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
        Mr, Mc, Gram, idx1, Vrhs = vie.jvie_system(Einc, e_r, r, freq, form=2)
        op = lib.Operator(L, M, N, precisions=lib.PREC_F64 | lib.PREC_F32)
        op.assemble_N(dx, em['ko'])
        op.set_material(Mr, Mc, Gram)
        op.set_index(idx1, one_based=True)
        cfg, _ = op.autotune(lib.PREC_F64, iters=5)
        op.autotune(lib.PREC_F32, iters=5)
        x = torch.randn(idx1.size, dtype=torch.complex128, device=dev)
        y = torch.empty_like(x)
        for _ in range(20):
            op.apply_device(x.data_ptr(), y.data_ptr(), lib.OP_AN, False, lib.PREC_F64)
        torch.cuda.synchronize()
        k, t = 200, time.perf_counter()
        for _ in range(k):
            op.apply_device(x.data_ptr(), y.data_ptr(), lib.OP_AN, False, lib.PREC_F64)
        torch.cuda.synchronize()
        gpu_mv = k / (time.perf_counter() - t)
        b = torch.from_numpy(np.ascontiguousarray(Vrhs)).to(dev)
        xs = torch.zeros_like(b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        flag, relres, it, _ = op.gmres_device(b.data_ptr(), xs.data_ptr(), restart=50, tol=1e-6, maxit=200)
        torch.cuda.synchronize()
        gpu_solve = time.perf_counter() - t
        iters = (int(it[1]) - 1) * 50 + int(it[0])
        row = {'workload': name, 'dof': int(idx1.size), 'tiles_fp64': cfg, 'gpu_matvecs_per_s': gpu_mv,
               'gpu_time_to_1e-6_s': gpu_solve, 'gpu_iterations': iters, 'gpu_flag': int(flag),
               'gpu_relres': float(relres)}
        # the refined solve (GMRES mode 3, cycles of 128), timed after one warm run
        for rep in range(2):
            x3 = torch.zeros_like(b)
            torch.cuda.synchronize()
            t = time.perf_counter()
            flag3, relres3, it3, rv3 = op.gmres_device(b.data_ptr(), x3.data_ptr(), restart=128, tol=1e-6,
                                                       maxit=200, mode=3, inner_tol=1e-7)
            torch.cuda.synchronize()
            t3 = time.perf_counter() - t
        row.update({'gpu_refined_mode3_time_to_1e-6_s': t3, 'gpu_refined_mode3_products': len(rv3) - 1 - it3[1],
                    'gpu_refined_mode3_flag': int(flag3), 'gpu_refined_mode3_relres': float(relres3),
                    'gpu_refined_mode3_rel_diff_vs_fp64': (torch.linalg.norm(x3 - xs) / torch.linalg.norm(xs)).item()})
        log('%s: refined mode 3 %.3f s (%d products), relres %.2e' % (name, t3, len(rv3) - 1 - it3[1], relres3))
        if not cpu:
            log('%s: GPU %.1f matvecs/s, solve %.3f s (%d it)' % (name, gpu_mv, gpu_solve, iters))
            rows.append(row)
            op.close()
            continue
        # CPU restatement on the same operator
        fN = op.get_fN()
        xh = x.cpu().numpy()
        idx0 = idx1 - 1
        omvp.mv_AN(xh, fN, Mr, Mc, Gram, 'notransp', idx0, workers=cores)
        reps, t = 0, time.perf_counter()
        while reps < 50 and time.perf_counter() - t < 10.0:
            omvp.mv_AN(xh, fN, Mr, Mc, Gram, 'notransp', idx0, workers=cores)
            reps += 1
        t_mv = (time.perf_counter() - t) / reps
        row.update({'cpu_matvecs_per_s': 1.0 / t_mv, 'cpu_cores': cores, 'cpu_matvec_sample': reps})
        row['cpu_model'] = cpu_model()
        if iters * t_mv < 200.0:
            calls = [0]

            def A(v):
                calls[0] += 1
                if calls[0] % 20 == 0:
                    log('  CPU solve: %d products' % calls[0])
                return omvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx0, workers=cores)
            t = time.perf_counter()
            xc, cflag, crel, cit, _ = okr.pgmres(A, Vrhs, restart=50, tol=1e-6, maxit=200)
            row.update({'cpu_time_to_1e-6_s': time.perf_counter() - t, 'cpu_flag': int(cflag),
                        'cpu_iterations': (int(cit[1]) - 1) * 50 + int(cit[0]),
                        'gpu_vs_cpu_solution_rel_diff': float(np.linalg.norm(xs.cpu().numpy() - xc)
                                                              / np.linalg.norm(xc))})
        else:
            row['cpu_time_to_1e-6_s_estimate'] = iters * t_mv
        log('%s: GPU %.1f matvecs/s, solve %.3f s (%d it); CPU %.2f matvecs/s, solve %s' % (
            name, gpu_mv, gpu_solve, iters, 1.0 / t_mv,
            ('%.1f s' % row['cpu_time_to_1e-6_s']) if 'cpu_time_to_1e-6_s' in row
            else 'est. %.0f s' % row['cpu_time_to_1e-6_s_estimate']))
        rows.append(row)
        op.close()
    return {'metric': 'other workloads', 'workloads': rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--size', type=int, default=128)
    ap.add_argument('--no-solve', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-block', action='store_true', help='skip the 16-excitation block solve')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the adjoint gradient timing')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the several-bodies batch (4 phantoms x 2 frequencies x 16 coil ports)')
    ap.add_argument('--profile-only', action='store_true',
                    help='assemble and run the timed FP64 matvecs only (for rocprofv3 runs)')
    ap.add_argument('--profile-solve', action='store_true',
                    help='assemble, then one FP64 and one FP32-refined solve to 1e-6 (for rocprofv3 runs)')
    ap.add_argument('--configs', action='store_true',
                    help='the other workloads instead: 32^3 sphere and 96x96x112 phantom at 128/298 MHz, '
                         'GPU and CPU-restatement matvecs/s and time to 1e-6')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='run a plain 1 GiB device copy five times (scale factor for FETCH/WRITE_SIZE)')
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 26, dtype=torch.complex128, device='cuda:0')   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': a.numel() * 16, 'copies': 5}))
        return
    if args.gpus != 1:
        print('bench.py measures one GPU: the package runs one handle per GPU and has no multi-GPU path; '
              'run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    import torch
    from marie_amd import geometry, lib, vie

    dev = torch.device('cuda:0')
    torch.cuda.init()
    if args.configs:
        print(json.dumps(other_workloads(torch, dev, geometry, lib, vie, cpu=not args.no_cpu)))
        return
    n1 = args.size
    L = M = N = n1
    dx = 1.5e-3 * 128 / n1
    freq = 298e6
    scale = 0.85
    r, eps_r, sig, ph = geometry.head_phantom(L, M, N, dx, seed=0, scale=scale)
    em = vie.EMconstants(freq)
    e_r = eps_r - 1j * sig / (em['eo'] * em['omega'])
    Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    Mr, Mc, Gram, idx1, Vrhs = vie.jvie_system(Einc, e_r, r, freq, form=2)
    nidx = idx1.size
    log('grid %d^3, dx %.3g mm, non-air DOF %d' % (n1, dx * 1e3, nidx))

    op = lib.Operator(L, M, N, precisions=lib.PREC_F64 | lib.PREC_F32)
    stream = torch.cuda.current_stream(dev)
    op.set_stream(stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.assemble_N(dx, em['ko'])
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    log('assembly %.3f s' % t_asm)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)

    g = torch.Generator(device='cpu').manual_seed(1234)
    xr = torch.randn(nidx, generator=g, dtype=torch.float64)
    xi = torch.randn(nidx, generator=g, dtype=torch.float64)
    x64 = torch.complex(xr, xi).to(dev)
    y64 = torch.empty_like(x64)
    x32 = x64.to(torch.complex64)
    y32 = torch.empty_like(x32)

    def run(prec, k):
        xp, yp = (x64, y64) if prec == lib.PREC_F64 else (x32, y32)
        for _ in range(k):
            op.apply_device(xp.data_ptr(), yp.data_ptr(), lib.OP_AN, False, prec)

    def timed(prec, k, w):
        run(prec, w)
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(prec, k)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    # tile autotuning (part of the set-up, like an FFT plan)
    tiles = {}
    for prec, name in ((lib.PREC_F64, 'fp64'), (lib.PREC_F32, 'fp32')):
        cfg, ms = op.autotune(prec, iters=5)
        tiles[name] = cfg
        log('tiles %s %s  (%s ms)' % (name, cfg, ' '.join('%.3f' % m for m in ms)))

    if args.profile_solve:
        b = torch.from_numpy(np.ascontiguousarray(Vrhs)).to(dev)
        for mode, rs in ((0, 50), (3, 128)):
            xs = torch.zeros_like(b)
            torch.cuda.synchronize()
            t = time.perf_counter()
            flag, relres, it, rv = op.gmres_device(b.data_ptr(), xs.data_ptr(), restart=rs, tol=1e-6, maxit=200,
                                                   mode=mode, inner_tol=1e-7)
            torch.cuda.synchronize()
            log('mode %d solve %.3f s, %d resvec entries, iter %s, relres %.2e' % (
                mode, time.perf_counter() - t, len(rv) - 1, list(it), relres))
        return
    K, W = args.steps, args.warmup
    T64 = timed(lib.PREC_F64, K, W)
    mv64 = K / T64
    log('fp64 %.1f matvecs/s (%.3f ms)' % (mv64, 1e3 * T64 / K))
    if args.profile_only:
        print(json.dumps({'metric': 'matvecs_per_s_fp64', 'value': mv64, 'ms_per_step': 1e3 * T64 / K}))
        return
    T32 = timed(lib.PREC_F32, K, W)
    mv32 = K / T32
    log('fp32 %.1f matvecs/s (%.3f ms)' % (mv32, 1e3 * T32 / K))
    # host-buffer entry point (marie_vie_apply): x and y cross PCIe every call,
    # as in the reference's gpuArray path (mv_AN.m:51, :141); reported beside
    # `value`, never as it
    xh = x64.cpu().numpy()
    op.apply(xh)
    th = time.perf_counter()
    for _ in range(10):
        op.apply(xh)
    mvh = 10 / (time.perf_counter() - th)
    log('fp64 host-buffer (PCIe-inclusive) %.1f matvecs/s' % mvh)

    # per-kernel durations from HIP events on the launch stream
    Px, Py, Pz = op.P
    kern = {}
    for prec, s, name in ((lib.PREC_F64, 16, 'fp64'), (lib.PREC_F32, 8, 'fp32')):
        run(prec, 10)                   # the host-buffer loop above leaves the device idle between products
        op.enable_timing(True)
        run(prec, 100)
        ms, calls = op.timing()
        op.enable_timing(False)
        kb = kernel_bytes(L, M, N, Px, Py, Pz, nidx, s)
        kern[name] = {k: {'ms': m, 'GBps': kb[k] / (m * 1e-3) / 1e9, 'bytes': kb[k]}
                      for k, m in zip(kb.keys(), ms)}
    dom = max(kern['fp64'], key=lambda k: kern['fp64'][k]['ms'])
    d = kern['fp64'][dom]
    n = L * M * N
    nS = nidx // 3
    bmv = 16 * (8 * nS + 48 * n)
    # measured HBM bytes per launch of the dominant kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    # scaled by a 1 GiB copy), recorded for this workload in profiles/
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                         'jvie_matvec_128cubed_fp64_traffic.json')
    if n1 == 128 and os.path.exists(tpath):
        tj = json.load(open(tpath))
        if dom in tj['kernels']:
            traffic = tj['kernels'][dom]['fetch_bytes'] + tj['kernels'][dom]['write_bytes']
    roof = {'bound': 'hbm', 'kernel': dom, 'achieved': d['GBps'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': d['GBps'] / HBM_PEAK_GBS, 'traffic': traffic, 'algorithmic_bytes': d['bytes'],
            'matvec_algorithmic_bytes': bmv, 'matvec_achieved_GBps': bmv * mv64 / 1e9,
            'matvec_frac': bmv * mv64 / 1e9 / HBM_PEAK_GBS}

    # Gram-Schmidt kernels of the solve (multi-dot, combination): their
    # rocprofv3 per-launch averages and algorithmic bytes, recorded for this
    # workload in profiles/
    kpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                         'jvie_gmres_128cubed_krylov_kernels.json')
    if n1 == 128 and os.path.exists(kpath):
        roof['krylov_kernels'] = json.load(open(kpath))['kernels']

    out = {}
    if not args.no_solve:
        b = torch.from_numpy(np.ascontiguousarray(Vrhs)).to(dev)

        def solve(mode, tol, inner_tol=1e-7, restart=50):
            xs = torch.zeros_like(b)
            torch.cuda.synchronize()
            t = time.perf_counter()
            flag, relres, it, rv = op.gmres_device(b.data_ptr(), xs.data_ptr(), restart=restart, tol=tol, maxit=200,
                                                   mode=mode, inner_tol=inner_tol)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            # resvec holds one entry per inner iteration (product) plus, in the
            # refined modes, the true residual after each of the it[1] cycles
            nit = len(rv) - 1 - (it[1] if mode else 0)
            return xs, flag, relres, nit, dt, None, it

        # tightly converged FP64 solution: the reference for the error of each 1e-6 solve
        xref, _, rel_ref, it_ref, t_ref, _, _ = solve(0, 1e-10)
        log('fp64 reference solve to 1e-10: %.3f s, %d iterations, relres %.2e' % (t_ref, it_ref, rel_ref))
        nref = torch.linalg.norm(xref)
        res = {}
        # the refined solve: complex64 products in block-orthogonalised cycles
        # of up to 128 (mode 3); the classical Gram-Schmidt refinements (modes
        # 2 and 1, GMRES(50)) beside it
        refined_restart = 128
        for mode, name, rs in ((0, 'fp64', 50), (3, 'fp32_refined', refined_restart),
                               (2, 'fp32_refined_cgs', 50), (1, 'fp32_refined_fp64basis', 50)):
            solve(mode, 1e-6, restart=rs)        # warm: the workspace allocation is not part of the timing
            xs, flag, relres, nit, dt, _, itc = solve(mode, 1e-6, restart=rs)
            err = (torch.linalg.norm(xs - xref) / nref).item()
            res[name] = (xs, flag, relres, nit, dt, err, itc, rs)
            log('%s solve %.3f s flag %d relres %.3g products %d cycles %d error vs 1e-10 FP64 %.2e'
                % (name, dt, flag, relres, nit, itc[1], err))
        x64s = res['fp64'][0]
        # per-solve algorithmic bytes: products and Gram-Schmidt of every
        # Arnoldi step (cycles of `restart` steps, the last one shorter), plus
        # in the refined mode one FP64 residual product per cycle and at the end
        nS = nidx // 3
        bmv64, bmv32 = 16 * (8 * nS + 48 * L * M * N), 8 * (8 * nS + 48 * L * M * N)
        per_solve = {}
        for name, mode, s_ in (('fp64', 0, 16), ('fp32_refined', 3, 8), ('fp32_refined_cgs', 2, 8)):
            nit, cycles, rs = res[name][3], res[name][6], res[name][7]
            lens = [rs] * (cycles[1] - 1) + [nit - rs * (cycles[1] - 1)]
            if mode == 3:
                by = block_solve_bytes(lens, nidx, bmv32, s_) + cycles[1] * bmv64
            else:
                by = solve_bytes(lens, nidx, bmv32 if mode else bmv64, s_) + (cycles[1] if mode else 1) * bmv64
            per_solve[name] = {'algorithmic_bytes': by, 'cycle_lengths': lens,
                               'achieved_GBps': by / res[name][4] / 1e9,
                               'frac_of_hbm_peak': by / res[name][4] / 1e9 / HBM_PEAK_GBS}
        out = {
            'time_to_1e-6_fp64_s': res['fp64'][4],
            'time_to_1e-6_fp32_refined_s': res['fp32_refined'][4],
            'speedup_fp32_refined': res['fp64'][4] / res['fp32_refined'][4],
            'fp32_refined_method': 'GMRES mode 3: complex64 products in block-orthogonalised cycles of up to %d '
                                   '(blocks of 8: local Gram-Schmidt per product, block Gram-Schmidt + CholeskyQR2 '
                                   'per block), FP64 iterative refinement on the true residual' % refined_restart,
            'products_fp32_refined': res['fp32_refined'][3], 'cycles_fp32_refined': res['fp32_refined'][6][1],
            'time_to_1e-6_fp32_refined_cgs_s': res['fp32_refined_cgs'][4],
            'gmres_iters_fp32_refined_cgs': res['fp32_refined_cgs'][3],
            'time_to_1e-6_fp32_refined_fp64basis_s': res['fp32_refined_fp64basis'][4],
            'fp32_refined_rel_error_vs_fp64_1e-10': res['fp32_refined'][5],
            'fp64_rel_error_vs_fp64_1e-10': res['fp64'][5],
            'fp32_refined_rel_diff_vs_fp64_1e-6': (torch.linalg.norm(res['fp32_refined'][0] - x64s)
                                                   / torch.linalg.norm(x64s)).item(),
            'relres_fp64': res['fp64'][2], 'relres_fp32_refined': res['fp32_refined'][2],
            'flag_fp64': res['fp64'][1], 'flag_fp32_refined': res['fp32_refined'][1],
            'gmres_iters_fp64': res['fp64'][3], 'gmres_iters_fp32_refined': res['fp32_refined'][3],
            'fp64_reference_solve': {'tol': 1e-10, 'iters': it_ref, 'relres': rel_ref, 's': t_ref},
            'error_definition': 'relres = ||b - A x|| / ||b|| with the FP64 product (pgmres.m:175) must be <= 1e-6 '
                                'for both solves; the error is ||x - x_ref|| / ||x_ref|| against the FP64 GMRES '
                                'solution converged to 1e-10 (the FP64 solver matches the oracle pgmres to 1e-9), '
                                'reported for the refined and for the FP64 1e-6 solve: the refined solve meets '
                                'the 1e-6 residual bound, and its error stays at the FP64 1e-6 solve level',
            'per_solve_bytes': per_solve,
        }

    if not args.no_solve and not args.no_block:
        # 16 excitations of the same body (plane waves from seeded directions and
        # polarisations): one block GMRES vs 16 single solves, all to 1e-6; on a
        # stream of its own (HIP graphs cannot be captured on the null stream)
        s_blk = 16
        bstream = torch.cuda.Stream(device=dev)
        op.set_stream(bstream.cuda_stream)
        rng = np.random.default_rng(7)
        cols = []
        for q in range(s_blk):
            dvec = rng.standard_normal(3)
            dvec /= np.linalg.norm(dvec)
            pol = np.cross(dvec, rng.standard_normal(3))
            pol /= np.linalg.norm(pol)
            Eq, _ = vie.PlaneWave_Excitation(r, em['ko'] * dvec, em['omega_mu'], pol)
            cols.append(vie.jvie_system(Eq, e_r, r, freq, form=2)[4])
        Bd = torch.from_numpy(np.ascontiguousarray(np.stack(cols))).to(dev)
        torch.cuda.synchronize()
        t = time.perf_counter()
        Xseq = torch.zeros_like(Bd)
        torch.cuda.synchronize()      # the solves run on non-blocking streams
        for q in range(s_blk):
            op.gmres_device(Bd[q].data_ptr(), Xseq[q].data_ptr(), restart=50, tol=1e-6, maxit=200)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t
        blk = {'rhs': s_blk, 'sequential_fp64_s': t_seq}
        for mode in (0, 1):
            X = torch.zeros_like(Bd)
            torch.cuda.synchronize()
            t = time.perf_counter()
            fl, rr, it, rv = op.block_gmres_device(Bd.data_ptr(), X.data_ptr(), s_blk, restart=40, tol=1e-6,
                                                   maxit=100, mode=mode, inner_tol=1e-7, use_graph=True)
            torch.cuda.synchronize()
            tb = time.perf_counter() - t
            diff = max((torch.linalg.norm(X[q] - Xseq[q]) / torch.linalg.norm(Xseq[q])).item() for q in range(s_blk))
            key = 'block_fp64' if mode == 0 else 'block_fp32_products'
            blk[key] = {'s': tb, 'block_steps_last_cycle_and_cycles': it, 'max_relres': float(rr.max()),
                        'flags_max': int(fl.max()), 'max_rel_diff_vs_single_solves': diff}
            log('block gmres mode %d: %.3f s for %d excitations (sequential FP64 %.3f s), cycles %d, max relres %.2e'
                % (mode, tb, s_blk, t_seq, it[1], rr.max()))
            # the same solve with every block step on the device, one HIP graph
            # per step index replayed in every cycle (MARIE_BLOCK_DEVICE)
            X = torch.zeros_like(Bd)
            torch.cuda.synchronize()
            t = time.perf_counter()
            fl, rr, it, rv = op.block_gmres_device(Bd.data_ptr(), X.data_ptr(), s_blk, restart=40, tol=1e-6,
                                                   maxit=100, mode=mode, inner_tol=1e-7, use_graph=True,
                                                   device_steps=True)
            torch.cuda.synchronize()
            td = time.perf_counter() - t
            blk[key]['device_driven_graph_steps'] = {
                's': td, 'block_steps_last_cycle_and_cycles': it, 'max_relres': float(rr.max()),
                'flags_max': int(fl.max())}
            log('  device-driven graph steps: %.3f s, cycles %d, max relres %.2e' % (td, it[1], rr.max()))
        # the same 16 columns one at a time with the refined solver (GMRES mode
        # 3, cycles of 128): fewer bytes per product than any block variant
        X = torch.zeros_like(Bd)
        torch.cuda.synchronize()      # the solves run on non-blocking streams
        op.block_gmres_device(Bd[:1].data_ptr(), X[:1].data_ptr(), 1, restart=128, tol=1e-6, maxit=200, mode=3,
                              inner_tol=1e-7, columns=True)       # warm: workspace allocation
        X = torch.zeros_like(Bd)
        torch.cuda.synchronize()
        t = time.perf_counter()
        fl, rr, it, rv = op.block_gmres_device(Bd.data_ptr(), X.data_ptr(), s_blk, restart=128, tol=1e-6, maxit=200,
                                               mode=3, inner_tol=1e-7, columns=True)
        torch.cuda.synchronize()
        tc = time.perf_counter() - t
        diff = max((torch.linalg.norm(X[q] - Xseq[q]) / torch.linalg.norm(Xseq[q])).item() for q in range(s_blk))
        blk['columns_fp32_refined_mode3'] = {'s': tc, 'cycles_all_columns': it[1], 'max_relres': float(rr.max()),
                                             'flags_max': int(fl.max()), 'max_rel_diff_vs_single_solves': diff,
                                             'speedup_vs_sequential_fp64': t_seq / tc}
        log('16 columns one at a time, refined mode 3: %.3f s (%.2fx the sequential FP64 solves), max relres %.2e'
            % (tc, t_seq / tc, rr.max()))
        out['batching_16_excitations'] = blk
        torch.cuda.synchronize()
        op.set_stream(stream.cuda_stream)

    if not args.no_solve and not args.no_batch and n1 == 128:
        # several uneven bodies and frequencies: 64^3, 80^3, 96^3, 128^3 phantoms
        # over the same 0.192 m domain (seeds 1-4) at 128 and 298 MHz, each with
        # the 16 ports of a coil of 0.1 x 0.05 m loops in two rings of 8 on a
        # cylinder of radius 0.15 m and length 0.2 m; each problem is one
        # handle (device assembly, own autotuned tiles) solved by block GMRES,
        # first one after the other, then all at once (marie_vie_block_gmres_batch)
        from marie_amd import coil
        ports = coil.cylinder_loop_array(16, 0.15, 0.1, 0.05, seg_per_side=8, rings=2, length=0.2)
        probs = []
        t = time.perf_counter()
        for nb, seed in ((64, 1), (80, 2), (96, 3), (128, 4)):
            dxb = 0.192 / nb
            rb, epsb, sigb, _ = geometry.head_phantom(nb, nb, nb, dxb, seed=seed, scale=0.85)
            for fb in (128e6, 298e6):
                emb = vie.EMconstants(fb)
                e_rb = epsb - 1j * sigb / (emb['eo'] * emb['omega'])
                Es, _ = coil.CCL_Excitation(rb, emb['ko'], 1.0, ports)
                Mrb, Mcb, Gb, idxb, v0 = vie.jvie_system(Es[0], e_rb, rb, fb, form=2)
                # form II right-hand side Gram * j w eo (e_r - 1) / e_r * Einc on the
                # DOFs (JVIE_Solver.m:151-156, :242-244), the same factor for every port
                fac = np.tile((Gb * 1j * emb['omega'] * emb['eo'] * (e_rb - 1) / e_rb).ravel(order='F'), 3)[idxb - 1]
                cols = [fac * Ep.ravel(order='F')[idxb - 1] for Ep in Es]
                assert np.allclose(cols[0], v0, rtol=1e-12, atol=0)
                opb = lib.Operator(nb, nb, nb, precisions=lib.PREC_F64 | lib.PREC_F32)
                stb = torch.cuda.Stream(device=dev)       # each problem on its own stream
                opb.set_stream(stb.cuda_stream)
                opb.assemble_N(dxb, emb['ko'])
                opb.set_material(Mrb, Mcb, Gb)
                opb.set_index(idxb, one_based=True)
                opb.autotune(lib.PREC_F64, iters=3)
                opb.autotune(lib.PREC_F32, iters=3)
                Bb = torch.from_numpy(np.ascontiguousarray(np.stack(cols))).to(dev)
                probs.append((nb, fb, opb, Bb, stb))
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t
        kw = dict(restart=20, tol=1e-6, maxit=100, mode=0, inner_tol=1e-7, use_graph=True)
        seq = []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for nb, fb, opb, Bb, _ in probs:
            Xb = torch.zeros_like(Bb)
            torch.cuda.synchronize()      # the solves run on non-blocking streams
            t1 = time.perf_counter()
            fl, rr, it, rv = opb.block_gmres_device(Bb.data_ptr(), Xb.data_ptr(), 16, **kw)
            torch.cuda.synchronize()
            seq.append({'voxels': nb, 'MHz': fb / 1e6, 'dof': int(Bb.shape[1]), 's': time.perf_counter() - t1,
                        'cycles': it[1], 'max_relres': float(rr.max()), 'flags_max': int(fl.max())})
        t_seq = time.perf_counter() - t
        Xs = [torch.zeros_like(p[3]) for p in probs]
        torch.cuda.synchronize()
        t = time.perf_counter()
        fls, rrs, its = lib.block_gmres_batch([p[2] for p in probs], [p[3].data_ptr() for p in probs],
                                              [x.data_ptr() for x in Xs], [16] * len(probs), **kw)
        torch.cuda.synchronize()
        t_bat = time.perf_counter() - t
        # one host thread driving every problem's device-driven block steps
        # (one HIP graph per step index, FP64 and complex64 products)
        dev_batch = {}
        for mode in (0, 1):
            Xd = [torch.zeros_like(p[3]) for p in probs]
            torch.cuda.synchronize()
            t = time.perf_counter()
            fld, rrd, itd = lib.block_gmres_batch([p[2] for p in probs], [p[3].data_ptr() for p in probs],
                                                  [x.data_ptr() for x in Xd], [16] * len(probs),
                                                  **dict(kw, mode=mode, device_steps=True))
            torch.cuda.synchronize()
            dev_batch['fp64' if mode == 0 else 'fp32_products'] = {
                's': time.perf_counter() - t, 'max_relres': float(max(r.max() for r in rrd)),
                'flags_max': int(max(f.max() for f in fld)), 'steps_and_cycles': itd}
            del Xd
        # the reference's loop: every port of every problem one FP64 GMRES(50) after another
        torch.cuda.synchronize()
        t = time.perf_counter()
        for nb, fb, opb, Bb, _ in probs:
            Xb = torch.zeros_like(Bb)
            torch.cuda.synchronize()      # the solves run on non-blocking streams
            for q in range(Bb.shape[0]):
                opb.gmres_device(Bb[q].data_ptr(), Xb[q].data_ptr(), restart=50, tol=1e-6, maxit=200)
        torch.cuda.synchronize()
        t_ports = time.perf_counter() - t
        # the columns one at a time with the refined solver (mode 3), the
        # problems one after the other and then all at once
        kwc = dict(restart=128, tol=1e-6, maxit=200, mode=3, inner_tol=1e-7, columns=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for nb, fb, opb, Bb, _ in probs:
            Xb = torch.zeros_like(Bb)
            torch.cuda.synchronize()      # the solves run on non-blocking streams
            opb.block_gmres_device(Bb.data_ptr(), Xb.data_ptr(), 16, **kwc)
        torch.cuda.synchronize()
        t_seq3 = time.perf_counter() - t
        Xs3 = [torch.zeros_like(p[3]) for p in probs]
        torch.cuda.synchronize()
        t = time.perf_counter()
        fl3, rr3, it3 = lib.block_gmres_batch([p[2] for p in probs], [p[3].data_ptr() for p in probs],
                                              [x.data_ptr() for x in Xs3], [16] * len(probs), **kwc)
        torch.cuda.synchronize()
        t_bat3 = time.perf_counter() - t
        out['batching_bodies_frequencies'] = {
            'problems': seq, 'ports_per_problem': 16, 'setup_s': t_setup, 'sequential_block_s': t_seq,
            'concurrent_batch_s': t_bat, 'speedup': t_seq / t_bat,
            'batch_max_relres': float(max(r.max() for r in rrs)), 'batch_flags_max': int(max(f.max() for f in fls)),
            'per_port_fp64_sequential_s': t_ports,
            'columns_fp32_refined_mode3': {'sequential_s': t_seq3, 'concurrent_batch_s': t_bat3,
                                           'speedup_concurrent_vs_sequential': t_seq3 / t_bat3,
                                           'speedup_vs_per_port_fp64': t_ports / t_bat3,
                                           'max_relres': float(max(r.max() for r in rr3)),
                                           'flags_max': int(max(f.max() for f in fl3))},
            'speedup_block_batch_vs_per_port_fp64': t_ports / t_bat,
            'device_driven_graph_batch': dict(dev_batch, method=(
                'MARIE_BLOCK_DEVICE | MARIE_BLOCK_GRAPH: every block step (product, block Gram-Schmidt, CholeskyQR, '
                'Householder least-squares update) on the device, one HIP graph per step index, all problems '
                'driven from one host thread by their status events'))}
        log('batch of %d problems x 16 ports: block GMRES one after the other %.3f s, concurrent %.3f s (max relres '
            '%.2e; setup %.1f s); per-port FP64 GMRES %.3f s; refined mode-3 columns one after the other %.3f s, '
            'concurrent %.3f s (max relres %.2e)' % (len(probs), t_seq, t_bat, max(r.max() for r in rrs), t_setup,
                                                     t_ports, t_seq3, t_bat3, max(r.max() for r in rr3)))
        log('  device-driven graph batch: FP64 %.3f s, complex64 products %.3f s'
            % (dev_batch['fp64']['s'], dev_batch['fp32_products']['s']))
        for p in probs:
            p[2].close()
        del probs, Xs

    if not args.no_solve and not args.no_adjoint:
        # differentiation: the same 128^3 phantom driven by the 8 ports of a
        # two-ring loop coil (0.1 x 0.05 m loops on a cylinder of radius 0.15 m,
        # length 0.2 m) in circular polarisation; J = coefficient of variation
        # of |B1+| over a central ellipsoid + L2 misfit to a seeded target;
        # gradient w.r.t. eps_r and sigma of every body voxel by adjoint solves
        # (Krylov basis recomputed, not stored), checked by Richardson central
        # differences on the voxels of largest |dJ/deps_r| and |dJ/dsigma|
        from marie_amd import adjoint, coil
        ports = coil.cylinder_loop_array(8, 0.15, 0.1, 0.05, seg_per_side=8, rings=2, length=0.2)
        Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
        op.assemble_K(dx, em['ko'])
        body = np.abs(Mc) > 1e-12
        prob = adjoint.JVIEProblem(op, body, Es, freq, dx, tol=1e-10, Hinc=Hs)
        nb = prob.idx.size // 3
        wq = torch.tensor(np.exp(-2j * np.pi * np.arange(8) / 4), device=dev)
        pts = r.reshape(-1, 3, order='F')[prob.idx[:nb]]
        roi = torch.tensor(np.sum((pts / np.array([0.04, 0.035, 0.05])) ** 2, axis=1) < 1.0, device=dev)

        def b1(e_, s_):
            X = adjoint.jvie_solve(e_, s_, prob)
            return torch.sum(wq[:, None] * adjoint.b1_plus(adjoint.total_B(X, prob), prob), dim=0)

        def objective(e_, s_, tgt):
            bb = b1(e_, s_)
            a = bb.abs()[roi]
            return torch.std(a) / torch.mean(a) + torch.sum((bb - tgt).abs() ** 2) / torch.sum(tgt.abs() ** 2)

        ev = torch.tensor(adjoint.to_voxels(eps_r), device=dev, requires_grad=True)
        sv = torch.tensor(adjoint.to_voxels(sig), device=dev, requires_grad=True)
        rng = np.random.default_rng(3)
        with torch.no_grad():
            b0 = b1(ev, sv)
            target = b0 * (1 + 0.2 * torch.tensor(rng.standard_normal(b0.shape), device=dev))
        torch.cuda.synchronize()
        t = time.perf_counter()
        J = objective(ev, sv, target)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t
        t = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t
        Jv = float(J.detach())
        g64 = (ev.grad.clone(), sv.grad.clone())
        # the same gradient with the refined solver in both directions (GMRES
        # mode 3, cycles of 128, to the same 1e-10)
        prob.mode, prob.restart = 3, 128
        ev3 = ev.detach().clone().requires_grad_(True)
        sv3 = sv.detach().clone().requires_grad_(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        J3 = objective(ev3, sv3, target)
        torch.cuda.synchronize()
        t_fwd3 = time.perf_counter() - t
        t = time.perf_counter()
        J3.backward()
        torch.cuda.synchronize()
        t_bwd3 = time.perf_counter() - t
        refined = {'forward_solves_s': t_fwd3, 'adjoint_solves_and_gradient_s': t_bwd3,
                   'objective_rel_diff': abs(float(J3.detach()) - Jv) / abs(Jv),
                   'grad_eps_rel_diff_vs_fp64': (torch.linalg.norm(ev3.grad - g64[0]) / torch.linalg.norm(g64[0])).item(),
                   'grad_sigma_rel_diff_vs_fp64': (torch.linalg.norm(sv3.grad - g64[1])
                                                   / torch.linalg.norm(g64[1])).item()}
        log('B1+ gradient with refined mode-3 solves: forward %.2f s, backward %.2f s; gradient vs FP64 %.1e (eps_r), '
            '%.1e (sigma)' % (t_fwd3, t_bwd3, refined['grad_eps_rel_diff_vs_fp64'], refined['grad_sigma_rel_diff_vs_fp64']))
        prob.mode, prob.restart = 0, 50
        checks = []
        # the difference quotient differentiates the objective as the re-solves
        # compute it: at 1e-10 their truncation error shifts J by ~1e-10 of itself
        # in a way that is not smooth across the perturbation, which against a
        # 1 % step (J moves by ~3e-7 of itself at the most sensitive voxel) is a
        # bias of up to ~1e-2 in the quotient. Re-solving to 2e-12 and stepping
        # by 10 % (Richardson over h and h/2 leaves an O(h^4) truncation, ~4e-7
        # here) brings it to ~1e-6; the gradient itself agrees to 9 digits
        # between tol 1e-10 and 1e-12
        fd_tol = 2e-12
        prob.tol, prob.maxit = fd_tol, 400
        for which, grad, base in (('eps_r', ev.grad, ev.detach()), ('sigma', sv.grad, sv.detach())):
            v = int(torch.argmax(grad.abs()))
            h = 1e-1 * float(base[v])

            def quotient(step):
                vals = []
                for sgn in (1, -1):
                    p = base.clone()
                    p[v] += sgn * step
                    with torch.no_grad():
                        vals.append(float(objective(p, sv.detach(), target) if which == 'eps_r'
                                          else objective(ev.detach(), p, target)))
                return (vals[0] - vals[1]) / (2 * step)

            try:
                fd = (4 * quotient(h / 2) - quotient(h)) / 3
            except lib.MarieError as exc:     # a re-solve that stalls above fd_tol
                checks.append({'parameter': which, 'voxel': v, 'autograd': float(grad[v]), 'error': str(exc)})
                continue
            checks.append({'parameter': which, 'voxel': v, 'autograd': float(grad[v]), 'central_difference': fd,
                           'step': h, 'rel_diff': abs(fd - float(grad[v])) / abs(fd)})
        out['differentiation_b1plus'] = {
            'ports': 8, 'parameters': 2 * int(body.sum()), 'objective': Jv, 'tol': 1e-10, 'fd_tol': fd_tol,
            'forward_solves_s': t_fwd, 'adjoint_solves_and_gradient_s': t_bwd, 'finite_difference_checks': checks,
            'refined_mode3': refined}
        log('B1+ gradient (8 ports, %d parameters): forward %.2f s, backward %.2f s; FD rel diff %s'
            % (2 * int(body.sum()), t_fwd, t_bwd, ', '.join('%s %.1e' % (c['parameter'], c.get('rel_diff', np.nan))
                                                            for c in checks)))

    cpu = None
    if not args.no_cpu:
        import scipy.fft  # noqa: F401
        from oracle import mvp as omvp   # CPU baseline leg only
        cores = os.cpu_count() or 1
        cores = min(cores, 16)
        fN = op.get_fN()
        xh = x64.cpu().numpy()
        idx0 = idx1 - 1
        omvp.mv_AN(xh, fN, Mr, Mc, Gram, 'notransp', idx0, workers=cores)   # warm-up
        reps, t = 0, time.perf_counter()
        while True:
            omvp.mv_AN(xh, fN, Mr, Mc, Gram, 'notransp', idx0, workers=cores)
            reps += 1
            el = time.perf_counter() - t
            if el > 10.0 or reps >= 20:
                break
        cpu = {'value': reps / el, 'unit': 'matvecs/s', 'cores': cores, 'kind': 'port', 'model': cpu_model(),
               'sample': '%d FP64 matvecs of the same %d^3 operator (oracle.mvp.mv_AN: numpy + scipy.fft, '
                         'workers=%d), %.1f s' % (reps, n1, cores, el)}
        log('cpu %.3f matvecs/s' % cpu['value'])
        if not args.no_solve:
            # time to 1e-6 on the CPU-runnable workload (the 32^3 sphere), GPU beside CPU
            sph = other_workloads(torch, dev, geometry, lib, vie, cpu=True,
                                  names=('sphere_32cubed_6p25mm_298MHz',))['workloads'][0]
            cpu['time_to_1e-6_sphere_32cubed'] = {
                k: sph[k] for k in ('cpu_time_to_1e-6_s', 'cpu_iterations', 'gpu_time_to_1e-6_s', 'gpu_iterations',
                                    'cpu_matvecs_per_s', 'gpu_matvecs_per_s', 'gpu_vs_cpu_solution_rel_diff')
                if k in sph}

    line = {
        'metric': 'VIE matvecs/s + time-to-1e-6 residual at 128^3 voxels, FP64 vs FP32-refined',
        'value': mv64, 'unit': 'matvecs/s', 'n_gpus': 1, 'steps': K, 'warmup': W,
        'ms_per_step': 1e3 * T64 / K, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'complex128', 'data': 'synthetic',
        'config': {'workload': 'jvie_matvec_head_phantom_%dcubed_1p5mm_298MHz' % n1 if n1 == 128 else
                   'jvie_matvec_head_phantom_%dcubed' % n1,
                   'grid': [L, M, N], 'circulant': [Px, Py, Pz], 'dx_m': dx, 'freq_Hz': freq,
                   'phantom_scale': scale, 'phantom': ph, 'non_air_dof': nidx, 'form': 2},
        'fp32_matvecs_per_s': mv32, 'fp32_ms_per_step': 1e3 * T32 / K,
        'fp64_host_buffer_matvecs_per_s': mvh,
        'assembly_s': t_asm,
        'tiles_elements_per_thread': tiles,
        'kernels': kern,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    line.update(out)
    print(json.dumps(line))


if __name__ == '__main__':
    main()

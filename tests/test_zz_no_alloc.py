"""marie_vie_apply_device_noalloc against marie_vie_apply_device.

The allocation-free entry point runs the same five kernels with their work
buffers in the library's per-device scratch array. Through the C-ABI, two
calls in a row (the second on a smaller grid with fewer right-hand sides, so
the array holds the first call's larger contents) must give exactly the
products of marie_vie_apply_device, for N/AN/K, both directions and both
precisions; a product larger than the array is refused.
"""
import numpy as np
import pytest

from oracle import operators as oop

pytestmark = pytest.mark.gpu


def _sym_fN(L, M, N, rng):
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    return oop.fft_operator(oop.circulant_nop(G))


def _problem(lib, L, M, N, seed):
    rng = np.random.default_rng(seed)
    op = lib.Operator(L, M, N, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(_sym_fN(L, M, N, rng))
    ko, dx = 2 * np.pi / 299792458.0 * 298e6, 6.25e-3
    op.assemble_K(dx, ko)
    Mr = 1 + rng.random((L, M, N)) + 1j * rng.random((L, M, N))
    Mc = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    op.set_material(Mr, Mc, 0.7)
    n3 = 3 * L * M * N
    idx = np.sort(rng.choice(n3, int(0.6 * n3), replace=False))
    op.set_index(idx)
    return op, idx.size, rng


def _compare(torch, lib, op, nidx, nrhs, rng):
    dev = torch.device('cuda:0')
    x = rng.standard_normal((nrhs, nidx)) + 1j * rng.standard_normal((nrhs, nidx))
    x64 = torch.from_numpy(x).to(dev)
    x32 = x64.to(torch.complex64)
    cases = [(lib.OP_AN, False), (lib.OP_AN, True), (lib.OP_N, False), (lib.OP_L, False), (lib.OP_K, False),
             (lib.OP_P, False)]
    for prec, xs in ((lib.PREC_F64, x64), (lib.PREC_F32, x32)):
        for opk, tr in cases:
            ya = torch.empty_like(xs)
            yb = torch.full_like(xs, float('nan'))
            torch.cuda.synchronize()
            op.apply_device(xs.data_ptr(), ya.data_ptr(), opk, tr, prec, nrhs)
            op.apply_device_noalloc(xs.data_ptr(), yb.data_ptr(), opk, tr, prec, nrhs)
            torch.cuda.synchronize()
            assert torch.equal(ya, yb), (prec, opk, tr, nrhs)


def test_noalloc_product_matches_apply_device():
    import torch
    from marie_amd import lib
    print("case allocation-free product: 20x18x16 with 2 right-hand sides, then 10x9x8 with 1", flush=True)
    assert lib.noalloc_scratch_bytes() == 128 << 20
    big, nb, rng = _problem(lib, 20, 18, 16, seed=1)
    _compare(torch, lib, big, nb, 2, rng)
    small, ns, rng2 = _problem(lib, 10, 9, 8, seed=2)
    _compare(torch, lib, small, ns, 1, rng2)
    # the preconditioner's material must exist before the allocation-free call
    fresh, nf, rng3 = _problem(lib, 10, 9, 8, seed=3)
    xf = torch.zeros(nf, dtype=torch.complex128, device='cuda:0')
    yf = torch.empty_like(xf)
    with pytest.raises(lib.MarieError):
        fresh.apply_device_noalloc(xf.data_ptr(), yf.data_ptr(), lib.OP_P)
    # a product whose work buffers exceed the array is refused before any launch
    huge = lib.Operator(96, 96, 96)      # 254 MB of work buffers in FP64
    with pytest.raises(lib.MarieError, match='exceed'):
        huge.apply_device_noalloc(xf.data_ptr(), yf.data_ptr())
    torch.cuda.synchronize()

"""CPU checks of the C-ABI library: it loads and exports every declared symbol.

No compute call is made (no GPU here); the product path refuses to run without
the library.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'marie_vie.h')


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(marie_vie_\w+)\s*\(', txt)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ('marie_vie_create', 'marie_vie_apply', 'marie_vie_apply_device', 'marie_vie_gmres',
              'marie_vie_gmres_device', 'marie_vie_assemble_N', 'marie_vie_set_fN', 'marie_vie_destroy'):
        assert n in names


def test_library_exports_every_declared_symbol():
    from marie_amd import lib
    path = lib.LIB_PATH
    if not os.path.exists(path):
        pytest.fail('libmarie_vie.so not built (run __graft_entry__.build())')
    so = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(so, n)]
    assert not missing, missing
    # the Python binding covers the same set
    bound = {n for n, _, _ in lib.SIGNATURES}
    assert bound == set(declared_functions())


def test_product_refuses_without_library(tmp_path):
    from marie_amd import lib
    with pytest.raises(lib.MarieError):
        lib._lib = None
        try:
            lib.load(str(tmp_path / 'missing.so'))
        finally:
            lib._lib = None


def test_vie_mirror_rejects_cpu_flag():
    from marie_amd import vie
    with pytest.raises(ValueError):
        vie.mv_AN([1.0], None, [[[1.0]]], [[[0.0]]], 1.0, 'notransp', [1], 0)


def test_device_pointers_carry_resolved_values():
    """Raw pointers handed to the library hold the tensor's values: a lazily
    conjugated view (torch.conj) shares the unconjugated memory even after
    .contiguous(), so the autograd harness resolves it first."""
    import torch
    from marie_amd import adjoint
    g = torch.tensor([1 + 2j, -3 + 0.5j], dtype=torch.complex128)
    c = adjoint._dense(torch.conj(g))
    assert not c.is_conj() and c.data_ptr() != g.data_ptr()
    raw = torch.view_as_real(c).numpy()
    assert np.array_equal(raw[:, 1], -g.imag.numpy())
    n = adjoint._dense(torch.conj(g).imag)
    assert np.array_equal(n.numpy(), -g.imag.numpy())


def test_circulant_side_rule():
    """2L when its prime factors are at most 13 (getOPERATORS.m's size), else
    the smallest even 7-smooth side >= 2L."""
    from marie_amd import lib
    assert [lib.circulant_side(n) for n in (8, 11, 39, 81, 96, 112, 128)] == [16, 22, 78, 162, 192, 224, 256]
    assert lib.circulant_side(17) == 36          # 34 = 2 * 17
    assert lib.circulant_side(102) == 210        # 204 = 4 * 3 * 17
    assert lib.circulant_side(133) == 270        # 266 = 2 * 7 * 19
    for n in range(1, 400):
        m = lib.circulant_side(n)
        assert m >= 2 * n and m % 2 == 0
        assert lib._largest_prime_factor(m) <= 13


def test_mex_gateways_call_only_declared_entry_points():
    """Every marie_vie_* function the MEX gateway sources call is declared in
    include/marie_vie.h (and so exported by the library)."""
    src = os.path.join(ROOT, 'marie_amd', 'mex')
    declared = set(declared_functions())
    used = set()
    for name in os.listdir(src):
        if name.endswith(('.cpp', '.hpp')):
            used |= set(re.findall(r'\b(marie_vie_\w+)\s*\(', open(os.path.join(src, name)).read()))
    assert used and used <= declared, used - declared


def test_mex_gateways_build_and_check_arguments_without_gpu():
    """The gateways compile against the MEX API emulation (tests/mex) and
    reject a wrong argument count as MATLAB errors before any device call."""
    import subprocess
    import sys
    subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'tests', 'mex')], check=True)
    sys.path.insert(0, os.path.join(ROOT, 'tests', 'mex'))
    import emul
    for name, nargs in (('mv_AN_hip', 3), ('jvie_gmres_hip', 4), ('mv_N_hip', 3), ('mv_L_hip', 4), ('mv_K_hip', 2),
                        ('jvie_bicgstab_hip', 5), ('getOPERATORS_hip', 1)):
        g = emul.Gateway(name)
        with pytest.raises(emul.MexError, match='MARIE:hip:nargin'):
            g(*([np.zeros(2)] * nargs))
        g('clear', nargout=0)

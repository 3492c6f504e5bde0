"""CPU restatement of MARIE's FFT-VIE matvec, operator assembly and Krylov solve.

THIS PACKAGE IS TEST INFRASTRUCTURE. It is the checker that the HIP library
(`marie_amd`) is compared against. Only `tests/`, `__graft_entry__.smoke()` and
the `cpu_baseline` leg of `bench.py` may import it; the product path
(`marie_amd`) never imports, links or executes anything under `oracle/`.

Every function cites the reference file:line it restates (paths relative to
the thanospol/MARIE tree). Arrays follow MATLAB's column-major semantics:
a voxel grid is a numpy array of shape (L, M, N[, 3]) and a "compact" vector
is `grid.ravel(order='F')[idx]` with 0-based `idx` (MATLAB's idx minus one).

Pinning (what checks each function, see tests/test_oracle_*.py):
  * mvp.mv_AN / mv_N / mv_L      -- brute-force dense Toeplitz sum on tiny grids,
                                    complex symmetry u^T N v = v^T N u, adjoint
                                    identity <A^H u, v> = <u, A v>, linearity.
  * operators.circulant_nop       -- parity invariants, zero middle planes,
                                    circulant == Toeplitz on a dense check.
  * operators.cubature_nop        -- oracle/_ref (reference C++ cubature_nop.cpp
                                    compiled from its own source) to 1e-14.
  * operators.gauss_1d            -- gauss_1d.m tables (30-digit values quoted in
                                    tests) and exactness on polynomials.
  * singular.*                    -- oracle/_ref DIRECTFN and DEMCEM builds; the
                                    closed-form static coincident-square
                                    integral; brute-force convergence.
  * krylov.pgmres                 -- dense numpy solve, Givens-GMRES equivalence,
                                    exact-arithmetic invariants of CGS+LS.
  * mie.mie_interior_field        -- closed-form small-sphere (Rayleigh) limit;
                                    the VIE solution on a sphere.
  * fields.plane_wave_excitation  -- closed-form cell average of exp(-jkz).
"""

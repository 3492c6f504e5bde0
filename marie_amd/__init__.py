"""MI355X-native FFT-VIE library for MARIE's JVIE hot path.

`marie_amd.lib` binds the C-ABI of libmarie_vie.so; `marie_amd.vie` mirrors
MARIE's MATLAB-facing functions (mv_AN, mv_N, mv_L, getOPERATORS, pgmres,
JVIE_Solver) on top of it. All compute runs in the HIP library on the GPU.
"""

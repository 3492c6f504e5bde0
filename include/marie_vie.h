/*
 * marie_vie.h -- C-ABI of the MI355X (gfx950) FFT-VIE library.
 *
 * Drop-in boundary for MARIE's JVIE hot path. The MATLAB-facing functions
 * keep their signatures: the MEX gateways marie_amd/mex/mv_AN_hip.cpp
 * (mv_AN.m's GPU branch) and marie_amd/mex/jvie_gmres_hip.cpp (pgmres on
 * the JVIE operator) forward to these entry points; every entry point below
 * names the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns an int status: 0 = OK, otherwise one of
 *    MARIE_ERR_*; marie_vie_last_error() gives a message (thread-local).
 *  - Complex arrays are interleaved (re, im) doubles, column-major exactly
 *    as MATLAB stores them: a voxel field is L x M x N (x fastest), a
 *    3-vector field L x M x N x 3, fN is Px x Py x Pz x 6.
 *  - "Host" entry points take caller-owned host buffers that are never
 *    retained; "_device" entry points take caller-owned device pointers
 *    (complex128 or complex64 per `prec`) on the handle's device.
 *  - A handle is reentrant across handles, not across concurrent calls on
 *    the same handle. All work runs on the handle's stream; host entry
 *    points synchronise before returning, device entry points do not.
 */
#ifndef MARIE_VIE_H
#define MARIE_VIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MARIE_OK 0
#define MARIE_ERR_ARG 1
#define MARIE_ERR_ALLOC 2
#define MARIE_ERR_HIP 3
#define MARIE_ERR_STATE 4
#define MARIE_ERR_SIZE 5
#define MARIE_ERR_NONFINITE 6

#define MARIE_PREC_F64 1 /* complex128 */
#define MARIE_PREC_F32 2 /* complex64  */

#define MARIE_OP_AN 0 /* Gram*Mr*x - Mc*N x   (mv_AN.m)  */
#define MARIE_OP_N 1  /* N x                  (mv_N.m)   */
#define MARIE_OP_L 2  /* N x - Gram*x         (mv_L.m)   */
#define MARIE_OP_K 3  /* K x                  (mv_K.m)   */
/* the JVIE preconditioner: mv_AN with the material (1./Mr, 1./Mr - 1)
 * (JVIE_Solver.m:213, :232, :278, :297); with Mr = 1 (form II) it is Gram*x */
#define MARIE_OP_P 4

typedef struct marie_vie_op marie_vie_op;

/* Library version (major*10000 + minor*100 + patch). */
int marie_vie_version(void);

/* Message of the last failing call on this thread. */
const char* marie_vie_last_error(void);

/* Create an operator for an L x M x N voxel grid whose circulant is
 * Px x Py x Pz (Px >= L, ...; getOPERATORS returns 2L x 2M x 2N,
 * src_vie/src_operators/getOPERATORS.m:82-90, circulant_nop.m:5).
 * `precisions` is a bitmask of MARIE_PREC_*. Replaces the device set-up
 * of JVIE_Solver.m:253-278 (gpuDevice, gpuArray upload). */
int marie_vie_create(marie_vie_op** h, int L, int M, int N, int Px, int Py, int Pz,
                     int precisions, int device);

/* Free every device buffer of the handle (JVIE_Solver.m:509-511 reset(dev)). */
int marie_vie_destroy(marie_vie_op* h);

/* Run the handle's work on an external stream (hipStream_t, NULL = own). */
int marie_vie_set_stream(marie_vie_op* h, void* stream);

/* Upload fN (Px x Py x Pz x 6 complex, host), the FFT of the circulant
 * N operator as returned by getOPERATORS(r,f,'N') (getOPERATORS.m:88-90).
 * The F32 copy, if enabled, is rounded from it. */
int marie_vie_set_fN(marie_vie_op* h, const double* fN);

/* Same from a device buffer of complex128. */
int marie_vie_set_fN_device(marie_vie_op* h, const void* fN_dev);

/* Download fN (complex128) to host. */
int marie_vie_get_fN(marie_vie_op* h, double* fN);

/* Assemble fN on the device for voxel size dx and wavenumber ko
 * (getOPERATORS(r,f,'N'), getOPERATORS.m:47-48, :74-96; assembly_nop.m,
 * circulant_nop.m, fft_operator.m). The reference's circulant is 2L x 2M x 2N;
 * any side >= 2L-1 embeds the same Toeplitz operator (the Toeplitz block at
 * the start, its reflection at the end, zeros between) and gives the same
 * product, which lets callers choose sides with small prime factors.
 * `singular` = 10 doubles (re,im of I_ST, I_EAc, I_EAo, I_VAc, I_VAo as
 * returned by singular_st/ea/va.m) or NULL for the library's own
 * machine-precision self-terms. */
int marie_vie_assemble_N(marie_vie_op* h, double dx, double ko, const double* singular);

/* getOPERATORS' other output types (getOPERATORS.m:88-95 for 'N', :108-115
 * for 'K'): type MARIE_TYPE_T returns the Toeplitz generator
 * (assembly_nop.m / assembly_kop.m: L x M x N x 6, or x 3 for K) and
 * MARIE_TYPE_C its circulant (circulant_nop.m / circulant_kop.m:
 * Px x Py x Pz x 6 or x 3), assembled on the device as marie_vie_assemble_N /
 * _K do, into the host buffer `out` (interleaved complex, column-major). The
 * handle's stored operator is not changed. op = MARIE_OP_N (with `singular`
 * as marie_vie_assemble_N) or MARIE_OP_K (`singular` ignored). */
#define MARIE_TYPE_T 0
#define MARIE_TYPE_C 1
int marie_vie_assemble_type(marie_vie_op* h, int op, double dx, double ko, const double* singular, int type,
                            double* out);

/* The five self-term integrals the library uses (10 doubles out). */
int marie_vie_singular_terms(double dx, double ko, double* out10);

/* The five self-terms as the reference computes them by default: method
 * 'DIRECTFN' of singular_st.m:15-18, singular_ea.m:17-28, singular_va.m:20-31
 * (getOPERATORS.m:40-42 default; the 20 triangle-pair integrals of
 * directfn_ws_st_const.cpp / directfn_ws_ea_const.cpp / directfn_ws_va_const.cpp
 * with the vertex orders of their MEX gateways), Np_1D = np Gauss points per
 * direction (assembly_nop.m:16 uses 20; 1 <= np <= 64). Evaluated on the
 * current device. Passing these values to marie_vie_assemble_N reproduces
 * getOPERATORS(r,f,'N') with its default self-term method. */
int marie_vie_singular_terms_directfn(double dx, double ko, int np, double* out10);

/* The K operator (curl of the vector potential: H = Hinc + K J / Gram,
 * H_field_Kop.m:69-76): fK is 2L x 2M x 2N x 3 complex, the FFT of the
 * circulant of getOPERATORS(r,f,'K') (getOPERATORS.m:101-115, assembly_kop.m,
 * circulant_kop.m, fft_operator.m), interleaved complex, column-major. Stored
 * beside fN; products with MARIE_OP_K (mv_K.m:55-95, no material, no
 * transposed form). Same octant storage rule as fN (component c odd under
 * reflection of axis c, gperiodic_coeff_kop.m). */
int marie_vie_set_fK(marie_vie_op* h, const double* fK);
int marie_vie_get_fK(marie_vie_op* h, double* fK);

/* Assemble fK on the device for voxel size dx and wavenumber ko > 0
 * (assembly_kop.m: the 12-pair cubature of cubature_kop.cpp with the 4x4 rule
 * for every offset and the 10x10 rule below n_medium = 5; the 21-pair 20x20
 * near rule of cubature_kop_near_sym.m with nearby_correct_kop.m for the 8
 * nearest offsets; times dx^4/(j ko)^2/(4 pi); circulant_kop.m; FFT). */
int marie_vie_assemble_K(marie_vie_op* h, double dx, double ko);

/* Material: Mr, Mc (L x M x N complex, host) and the Gram scalar dx^3
 * (mv_AN.m:1 arguments Mr, Mc, Gram; JVIE_Solver.m:135-140, :242-244). */
int marie_vie_set_material(marie_vie_op* h, const double* Mr, const double* Mc, double Gram);

/* Compact DOF index into the L x M x N x 3 grid (idxS3, JVIE_Solver.m:147-148).
 * `one_based` = 1 for MATLAB indices. Duplicates are rejected. */
int marie_vie_set_index(marie_vie_op* h, const int64_t* idx, int64_t nidx, int one_based);

/* y = op(x) for `nrhs` stacked compact vectors of nidx complex each
 * (mv_AN.m:1 with transp_flag 'notransp'/'transp', mv_N.m, mv_L.m).
 * Host buffers, complex128. */
int marie_vie_apply(marie_vie_op* h, int op, int transp, const double* x, double* y, int nrhs);

/* Same on device buffers of precision `prec` (MARIE_PREC_F64 or _F32). */
int marie_vie_apply_device(marie_vie_op* h, int op, int transp, int prec, const void* x_dev,
                           void* y_dev, int nrhs);

/* marie_vie_apply_device without any allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). The
 * product's work buffers live in one file-scope device array per GPU of
 * marie_vie_noalloc_scratch_bytes() bytes (128 MiB), held until the process
 * ends; a product whose buffers, 16 (complex128) or 8 (complex64) bytes x
 * 3 nrhs Px N (M + Py) plus alignment, exceed it returns MARIE_ERR_SIZE.
 * Same kernels and results as marie_vie_apply_device; no timing events.
 * MARIE_OP_P needs Mr = 1 (form II) or its material derived by an earlier
 * MARIE_OP_P product of marie_vie_apply_device, else MARIE_ERR_STATE. Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU race
 * (serialise them, e.g. on one stream). The array needs no clearing: each
 * product writes every element it reads. */
int marie_vie_apply_device_noalloc(marie_vie_op* h, int op, int transp, int prec, const void* x_dev,
                                   void* y_dev, int nrhs);
size_t marie_vie_noalloc_scratch_bytes(void);

/* Restarted GMRES(restart), 1 <= restart <= 511 (<= 503 in mode 3), on A_N with MARIE's pgmres semantics
 * (src_numeric/src_iterative/pgmres.m:70-181, pgmres_iter.m:22-150):
 * stop when the least-squares residual < tol*||b||; flag/relres from the
 * true residual ||b - A x||/||b||; resvec relative to ||b|| (capacity
 * `resvec_cap`, count in *nres); iter = {inner, outer}. Device buffers
 * b_dev, x_dev (complex128, nidx each; x_dev holds x0 on entry).
 * `mode`: 0 = plain FP64; 1 = complex64 matvecs inside FP64 GMRES (FP64
 * Krylov basis) with iterative refinement on the FP64 true residual; 2 = as 1
 * with the Krylov basis stored in complex64 (dot products, norms, Givens
 * rotations and the refinement residual still in FP64). In modes 1 and 2 each
 * refinement step is one GMRES(restart) cycle on the FP64 residual r, stopped
 * when its least-squares residual reaches max(inner_tol*||r||,
 * 0.9*tol*||b||) (at most 0.5*||r||); x and r = b - A x are then updated in
 * FP64, so the residual is replaced by the true one at every restart.
 * 3 = the refinement of mode 2 with communication-avoiding (s-step) cycles:
 * up to 8 complex64 products of a scaled power basis back to back, then one
 * block Gram-Schmidt pass and CholeskyQR2 of the block on the device
 * (complex64 basis, FP64 accumulation), the Hessenberg columns and Givens
 * rotations on the host; at most `restart` products per refinement step and
 * the same stopping rule. No counterpart in the reference (its pgmres is
 * mode 0); iter[0] counts the last cycle's accepted products. */
int marie_vie_gmres_device(marie_vie_op* h, const void* b_dev, void* x_dev, int restart, double tol,
                           int maxit, int mode, double inner_tol, int* flag, double* relres,
                           int* iter, double* resvec, int resvec_cap, int* nres);

/* Same with host buffers b, x (complex128, nidx each; x holds x0 on entry
 * and the solution on return) -- the form a MEX gateway of pgmres /
 * JVIE_Solver.m:363-367 forwards to. */
int marie_vie_gmres(marie_vie_op* h, const double* b, double* x, int restart, double tol, int maxit,
                    int mode, double inner_tol, int* flag, double* relres, int* iter, double* resvec,
                    int resvec_cap, int* nres);

/* Preconditioned GMRES(restart), FP64, as JVIE_Solver.m:327, :347 call
 * pgmres with the preconditioner handle fPGPU (mv_AN with 1./Mr, 1./Mr - 1;
 * MARIE_OP_P): side MARIE_PRECOND_RIGHT solves A P z = b and returns
 * x = P z (pgmres R1, pgmres_iter.m:64-76, :131-145); MARIE_PRECOND_LEFT
 * solves P A x = P b with the stopping test and resvec against ||P b||
 * (pgmres L1, pgmres.m:122-142, pgmres_iter.m:80-96). flag and relres from
 * the true residual ||b - A x||/||b|| (pgmres.m:175-180). Other arguments as
 * marie_vie_gmres_device with mode 0. */
#define MARIE_PRECOND_NONE 0
#define MARIE_PRECOND_RIGHT 1
#define MARIE_PRECOND_LEFT 2
int marie_vie_gmres_precond_device(marie_vie_op* h, const void* b_dev, void* x_dev, int restart, double tol,
                                   int maxit, int side, int* flag, double* relres, int* iter, double* resvec,
                                   int resvec_cap, int* nres);

/* Same with host buffers b, x (complex128, nidx each). */
int marie_vie_gmres_precond(marie_vie_op* h, const double* b, double* x, int restart, double tol, int maxit,
                            int side, int* flag, double* relres, int* iter, double* resvec, int resvec_cap,
                            int* nres);

/* GMRES with deflated restarts, FP64: JVIE_Solver.m:378-450 (solver 'D')
 * calling pgmresDR(afun, Vrhs, inner_it, tol, outer_it, ritz[, L1 | R1])
 * (src_numeric/src_iterative/pgmresDR.m:73-282, pgmresDR_iter.m:23-142,
 * getHarmVecs1.m, getHarmVecs2.m): a GCRO-DR cycle that keeps the `ritz`
 * harmonic Ritz vectors of smallest magnitude (U, C = A U) between restarts.
 * ritz < 1 selects min(10, round(restart/2)); larger values are limited to
 * restart - 1 and to 32. side as marie_vie_gmres_precond_device. iter =
 * {steps of the last cycle, cycles after the first}; resvec relative to
 * ||b|| (||P b|| for left preconditioning); flag/relres by the true
 * residual. restart 2..64. */
int marie_vie_gmres_dr_device(marie_vie_op* h, const void* b_dev, void* x_dev, int restart, double tol,
                              int maxit, int ritz, int side, int* flag, double* relres, int* iter,
                              double* resvec, int resvec_cap, int* nres);

/* Same with host buffers b, x (complex128, nidx each). */
int marie_vie_gmres_dr(marie_vie_op* h, const double* b, double* x, int restart, double tol, int maxit,
                       int ritz, int side, int* flag, double* relres, int* iter, double* resvec,
                       int resvec_cap, int* nres);

/* BiCGStab, FP64, the solver JVIE_Solver.m:463-490 calls (MATLAB's
 * bicgstab(afun, Vrhs, tol, max_it[, pfun]) with pfun = MARIE_OP_P when
 * `precond` = 1, applied to p and s as a function-handle M). Van der Vorst's
 * recurrences with shadow residual r0; resvec: absolute residual norm of
 * every half step, starting with ||b - A x0||; stops when a half- or
 * full-step norm <= tol*||b||; *iter = half steps / 2 (k - 0.5 when it stops
 * halfway); flag 0 converged, 1 maxit reached, 4 breakdown (a zero <r~,r>,
 * <r~,v>, <t,t> or omega), by the true residual. x_dev holds x0 on entry.
 * resvec needs up to 2*maxit + 1 entries (count in *nres). */
int marie_vie_bicgstab_device(marie_vie_op* h, const void* b_dev, void* x_dev, double tol, int maxit,
                              int precond, int* flag, double* relres, double* iter, double* resvec,
                              int resvec_cap, int* nres);

/* Same with host buffers b, x (complex128, nidx each). */
int marie_vie_bicgstab(marie_vie_op* h, const double* b, double* x, double tol, int maxit, int precond,
                       int* flag, double* relres, double* iter, double* resvec, int resvec_cap, int* nres);

/* GMRES on the adjoint system A^H y = g (A^H applied through mv_AN's
 * 'transp' branch, mv_AN.m:65-98, the product MATLAB's solvers request as
 * afun(x,'transp'), iterapp.m:42-58). Arguments and semantics as
 * marie_vie_gmres_device. */
int marie_vie_gmres_adjoint_device(marie_vie_op* h, const void* g_dev, void* y_dev, int restart, double tol,
                                   int maxit, int mode, double inner_tol, int* flag, double* relres,
                                   int* iter, double* resvec, int resvec_cap, int* nres);

/* Per-voxel gradient of a real objective J(x) with respect to eps_r and
 * sigma (e_r = eps_r - j sigma/(eo omega), Scat_Solver.m:144) for the JVIE
 * form II system (JVIE_Solver.m:131-156, :242-244: Mr = 1,
 * Mc = (e_r-1)/e_r, b = j omega eo Gram Mc Einc) loaded in the handle.
 * x_dev: the solution; einc_dev: Einc on the DOF index (nidx complex128);
 * lambda_dev: the adjoint state, A^H lambda = dJ/dconj(x); omega_eo = omega*eo.
 * Writes grad_eps_dev, grad_sigma_dev (device, L*M*N doubles each, voxel
 * order i + L(j + M k)): dJ/deps_r = 2 Re S_v, dJ/dsigma = 2 Im S_v/(omega eo)
 * with S_v = sum_c conj(lambda_vc) (1 - Mc_v)^2 (N x + j omega eo Gram Einc)_vc;
 * voxels outside the DOF index get 0. No counterpart in the reference (it has
 * no derivatives); this is the adjoint of its forward solve. */
int marie_vie_material_gradient_device(marie_vie_op* h, const void* lambda_dev, const void* x_dev,
                                       const void* einc_dev, double omega_eo, double* grad_eps_dev,
                                       double* grad_sigma_dev);

/* Block GMRES for `nrhs` right-hand sides of the operator (device buffers B,
 * X: nrhs stacked compact vectors of nidx complex128 each; X holds X0 on
 * entry). The reference solves excitation ports one pgmres at a time
 * (VSIE_Solver.m:387-407, fVIEsolver_Assembly.m:185-197); here one block
 * Krylov space serves all columns: restarted block GMRES with `restart`
 * block steps per cycle, block classical Gram-Schmidt, CholeskyQR2 (shifted
 * CholeskyQR3 for an ill-conditioned block)
 * for the block QR, a Householder least-squares update per step, and the
 * FP64 true residual B - A X at every restart. Column q converges when
 * ||b_q - A x_q|| <= tol ||b_q|| (flag[q], relres[q] as pgmres.m:175).
 * mode 0 = complex128 products; mode 1 = complex64 products, FP64 basis and
 * FP64 residual replacement (each cycle stops at max(inner_tol*||r_q||,
 * 0.9*tol*||b_q||) per column). `options` is a bitmask:
 * MARIE_BLOCK_GRAPH replays each block step's multi-column product from a
 * HIP graph captured on the cycle that first reaches it (not on the legacy
 * default stream); MARIE_BLOCK_CGS2 orthogonalises with two classical
 * Gram-Schmidt passes instead of the one of pgmres_iter.m:106-107 (twice the
 * block-vector traffic, orthogonality kept to working precision). iter =
 * {block steps of the
 * last cycle, cycles}; resvec gets the worst column's relative residual per
 * step and, after each cycle, its true value. Limits: nrhs <= 64,
 * (restart+1)*nrhs <= 2048 basis vectors.
 * MARIE_BLOCK_COLUMNS solves the columns one at a time instead, each as
 * marie_vie_gmres_device in `mode` (0-3; 3: block-orthogonalised complex64
 * cycles, the fastest per column at 128^3) with `restart` its cycle length
 * (then no nrhs limit); iter = {largest last-cycle count, cycles of all
 * columns}, resvec each column's final relres.
 * MARIE_BLOCK_DEVICE (modes 0-1, nrhs <= 32) runs every part of a block step
 * on the device -- the product, the Gram-Schmidt pass, CholeskyQR (the s x s
 * factor on the device, the third pass skipped on the device when the first
 * factors unshifted) and the Householder least-squares update -- so a step
 * is one stream-ordered sequence without host round trips; with
 * MARIE_BLOCK_GRAPH the whole step is captured into a HIP graph once per step
 * index and replayed in every cycle. The host reads one status record per
 * step (the columns' least-squares residuals) and one per cycle (the true
 * residual norms); the arithmetic and the stopping rule are those above. */
#define MARIE_BLOCK_GRAPH 1
#define MARIE_BLOCK_CGS2 2
#define MARIE_BLOCK_COLUMNS 4
#define MARIE_BLOCK_DEVICE 8
int marie_vie_block_gmres_device(marie_vie_op* h, const void* B_dev, void* X_dev, int nrhs, int restart,
                                 double tol, int maxit, int mode, double inner_tol, int options,
                                 int* flag, double* relres, int* iter, double* resvec, int resvec_cap,
                                 int* nres);

/* Same with host buffers B, X (nrhs stacked compact vectors, complex128). */
int marie_vie_block_gmres(marie_vie_op* h, const double* B, double* X, int nrhs, int restart, double tol,
                          int maxit, int mode, double inner_tol, int options, int* flag, double* relres,
                          int* iter, double* resvec, int resvec_cap, int* nres);

/* Several problems (bodies, frequencies: separate handles), each with its
 * block of right-hand sides, solved concurrently: one host worker per
 * problem (at most 8 at a time, largest first) runs
 * marie_vie_block_gmres_device on that handle, whose stream is its own, so
 * the products and Gram-Schmidt kernels of different problems overlap on
 * the device. ops[i], B[i], X[i], nrhs[i] as the single-problem entry;
 * restart, tol, maxit, mode, inner_tol, options shared. flag and relres
 * hold sum(nrhs) entries in problem order, iter 2 per problem, status the
 * return code of each problem. With MARIE_BLOCK_DEVICE one host thread drives
 * every problem instead: each problem's block steps (HIP graphs with
 * MARIE_BLOCK_GRAPH) are launched on its stream, and whichever problem's step
 * has completed -- its status event -- gets its next step, its cycle's end or
 * its next cycle enqueued, so the problems' steps overlap on the device
 * without host workers. The batch form of fVIEsolver_Assembly's per-port
 * solves over several bodies (VSIE_Solver.m:387-407 per body). */
int marie_vie_block_gmres_batch(marie_vie_op* const* ops, int nitems, const void* const* B_dev,
                                void* const* X_dev, const int* nrhs, int restart, double tol, int maxit,
                                int mode, double inner_tol, int options, int* flag, double* relres,
                                int* iter, int* status);

/* Incident E and H of current elements at observation points (eval_DGF.m:37-106
 * as called by CCL_Excitation.m:41-70 and Dipole_Excitation.m): segment
 * midpoints seg_center (nseg x 3, row per segment), current elements seg_J
 * (nseg x 3 complex interleaved, current times segment vector), grouped into
 * ports by port_start[0..nport] (segments port_start[p] .. port_start[p+1]-1
 * drive port p); points column-major npoints x 3 (MATLAB's
 * reshape(r, L*M*N, 3)); ko = omega/c0, omega_mu = omega*mu0. Outputs E, H:
 * nport x 3 x npoints complex (per port the column-major npoints x 3 array of
 * reshape(E, L, M, N, 3)). No counterpart in the C of the reference (eval_DGF
 * is MATLAB); this is its device form. */
int marie_vie_dgf_fields(const double* seg_center, const double* seg_J, const int* port_start, int nport,
                         const double* points, int64_t npoints, double ko, double omega_mu, double* E,
                         double* H, int device);
/* Same on device buffers, on `stream` (synchronised before returning). */
int marie_vie_dgf_fields_device(const double* seg_center_dev, const void* seg_J_dev, const int* port_start_dev,
                                int nport, const double* points_dev, int64_t npoints, double ko, double omega_mu,
                                void* E_dev, void* H_dev, void* stream);

/* Tile autotuner: times every compiled variant of the five matvec kernels
 * with HIP events on the handle's stream, `iters` products each, picks the
 * fastest per kernel for precision `prec`, then revisits each kernel with
 * the others at their picks and keeps the variant that gives the fastest
 * whole product (a kernel's time depends on what its neighbours leave in the
 * caches). A variant is cfg6[0..4], one
 * value per kernel (fwd_x, fwd_y, z, inv_y, inv_x): v > 0 selects the
 * LDS-stage FFT with v elements per thread, v < 0 the register-resident FFT
 * with -v points per thread (power-of-two sides 32-512 with 4, 8 or 16
 * points; mixed-radix sides 96/192 with 12 or 24, 112/224 with 14 or 28, 160
 * with 10 or 20 -- z: 12, 10, and 14 in complex64); for the z kernel,
 * v = -(100*TX + E) selects the variant with one thread group per field
 * component (TX columns per workgroup, E points per thread, the three
 * spectra exchanged through LDS for the tensor product); for the y kernels,
 * v = -(200 + E) selects two lines per thread on power-of-two sides
 * (complex64: adjacent column pairs, one 16-byte access per point; FP64:
 * columns TX apart, each access a contiguous run); cfg6[5] is reserved and
 * returned as 0. cfg6 / best_ms5 (optional) receive the choice and its mean
 * kernel times. No counterpart in the reference (MATLAB's fftn plans
 * internally). */
int marie_vie_autotune(marie_vie_op* h, int prec, int iters, int* cfg6, double* best_ms5);

/* Set the kernel variants explicitly (cfg6 as above). A new handle starts at
 * {-4, -16, -8, -208, -4} for FP64 and {-8, -16, -16, -16, -8} for FP32 (the
 * autotuner's usual choice at 256-point sides); on a mixed-radix side a
 * register code takes that side's own plan (12, 14 or 20 points; the z
 * kernel at 224 one thread group per component, -414), and a kernel whose
 * side the register engine does not cover takes the LDS-stage FFT. */
int marie_vie_set_tiles(marie_vie_op* h, int prec, const int* cfg6);

/* Per-kernel timing with HIP events on the handle's stream. When enabled,
 * every apply records an event between its five kernels and accumulates
 * the durations; marie_vie_get_timing returns the mean ms per kernel
 * (5 values: fwd_x, fwd_y, z_product, inv_y, inv_x_epilogue) and the count. */
int marie_vie_enable_timing(marie_vie_op* h, int on);
int marie_vie_get_timing(marie_vie_op* h, double* mean_ms5, int64_t* calls);

#ifdef __cplusplus
}
#endif
#endif /* MARIE_VIE_H */

/* dgmi.h -- C ABI of libdgmi.so, the MI355X-native interior-penalty DG diffusion solve.
 *
 * The reference (linfeng810/dg_nn_public) has no FFI: its operator API is a set of
 * module-level Python functions over torch tensors that read global state
 * (z02/cmmn_data.py:19-64).  Each entry point below replaces one of those functions;
 * the Python layer in dg_nn_public_amd/ keeps the reference's names and signatures on
 * top of it.  Citations are relative to the reference repository root,
 * z02/ = z02-implicit/z06-DGtoCG/, z05/ = z05-ns/.
 *
 * Conventions
 *   - every entry point returns 0 on success and a negative code on error; the message
 *     is available from dgmi_last_error() (thread-local);
 *   - vectors are caller-owned device pointers in the reference's layout: flat
 *     (nele*nloc), element-major; the context owns geometry, tables, multigrid data and
 *     scratch;
 *   - `dtype` selects the arithmetic of an operator apply: DGMI_F64, DGMI_F32, or
 *     DGMI_BF16 (bf16 inputs, fp32 accumulation; vectors are then fp32);
 *   - `stream` is a hipStream_t passed as void* (0 = the null stream); no entry point
 *     synchronises the device unless it says so.
 *   - a context is not safe for concurrent use from several threads.
 */
#ifndef DGMI_H
#define DGMI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgmi_ctx dgmi_ctx;

enum { DGMI_F64 = 0, DGMI_F32 = 1, DGMI_BF16 = 2 };
enum { DGMI_SMOOTH_JACOBI3 = 0, DGMI_SMOOTH_DIRECT = 1 };

typedef struct {
  int32_t D;               /* 2 (P3 triangles) or 3 (P3 tetrahedra)                      */
  int64_t nele;            /* number of elements                                        */
  int64_t nverts;          /* number of P1 vertices (P1CG nodes)                         */
  const double* x_ref_in;  /* host (nele, D, nloc) P3 node coordinates                   */
  const int64_t* cells;    /* host (nele, D+1) vertex ids of each element                */
  const double* verts;     /* host (nverts, D) vertex coordinates                        */
  const double* kappa;     /* host (nele) diffusivity per element                        */
  const int64_t* bctype;   /* host (nele*(D+1)) face types, or NULL: boundary faces are
                              Dirichlet (0), matched faces interior (-1); Neumann faces
                              (1) carry no operator term and add their flux to the
                              right-hand side; other types carry no term at all
                              (z02/mesh_init.py:697-741, z02/volume_mf_diff.py:1575-1589) */
  double eta_e;            /* penalty coefficient (z02/config.py:245: 36)                 */
  int32_t face_kappa;      /* interior-face diffusivity: 0 = arithmetic mean, the
                              reference's (z05/volume_mf_um_diff_varprop.py:567);
                              1 = harmonic mean (symmetric weighted interior penalty,
                              coercive at any contrast)                                  */
  void* stream;            /* hipStream_t of the creation's device work and copies, or
                              NULL (the null stream); creation returns when it is done   */
} dgmi_mesh_desc;

/* Context: topology (z02/mesh_init.py:260-658), reference-element tables
 * (z02/shape_function.py:459-1317), affine geometry, uploaded to the current device. */
int dgmi_ctx_create(const dgmi_mesh_desc* desc, dgmi_ctx** out);
/* Gives the context's device memory back to the library's per-device block cache, from
 * which later contexts of about the same sizes are served without hipMalloc (hipFree waits
 * for the whole device to go idle).  Synchronises the device first unless the parameter
 * "destroy_sync" is 0, with which the caller promises that all work it queued for the
 * context has finished. */
int dgmi_ctx_destroy(dgmi_ctx* ctx);
/* Frees the idle blocks of the block cache on every device (they are also freed, and the
 * allocation retried, when a hipMalloc fails).  No counterpart in the reference. */
int dgmi_release_cached_memory(void);
/* A context holding ncomp independent solves ("components") packed into one mesh: the
 * union of their cells and vertices, component c owning elements [elem_off[c],
 * elem_off[c+1]) and vertices [vert_off[c], vert_off[c+1]).  Operator applies, smoothing
 * and every V-cycle level run as one launch over all components; each component keeps its
 * own SFC hierarchy and Krylov space, and dgmi_gmres_solve converges each to its own
 * tolerance (stats then holds 4 values per component).  No counterpart in the reference,
 * which solves one mesh per process; it is the packed-launch entry of the batch and
 * serving workloads. */
int dgmi_ctx_create_batch(const dgmi_mesh_desc* desc, int32_t ncomp, const int64_t* elem_off,
                          const int64_t* vert_off, dgmi_ctx** out);
/* dgmi_ctx_create_batch with component c's P3 node coordinates at x_parts[c] (host,
 * (nele_c, D, nloc)) instead of one concatenated desc->x_ref_in (ignored): the pieces are
 * copied to the device one by one, with no host copy of the union. */
int dgmi_ctx_create_batch_parts(const dgmi_mesh_desc* desc, int32_t ncomp,
                                const int64_t* elem_off, const int64_t* vert_off,
                                const double* const* x_parts, dgmi_ctx** out);
/* info[0..7] = nele, nloc, ndof, nface, ngi, sngi, nverts, device bytes held */
int dgmi_ctx_info(const dgmi_ctx* ctx, int64_t* info);
/* Host wall-clock seconds spent in the set-up phases of this context so far, the analogue
 * of the reference's time.time() prints around its set-up (z02/main.py:32-110):
 * seconds[0] topology, [1] geometry, [2] device tables and uploads (dgmi_ctx_create);
 * [3] block diagonal and smoother (dgmi_smoother_setup); [4] P1CG graph, colouring and SFC
 * order, [5] RAR probing (up to the host copy of RAR), [6] SFC level operators, [7] device
 * layout of the levels and the coarse inverses (dgmi_mg_setup). */
int dgmi_setup_timing(const dgmi_ctx* ctx, double* seconds);
/* host copies of nbf, alnmt, glb_bcface_type (each nele*nface) */
int dgmi_topology(const dgmi_ctx* ctx, int64_t* nbf, int64_t* alnmt, int64_t* bctype);
/* host copy of the element tables: n (nloc*ngi), nlx (D*nloc*ngi), weight (ngi),
 * sn (nface*nloc*sngi), snlx (nface*D*nloc*sngi), sweight (sngi), gi_align (D*sngi) */
int dgmi_element_tables(const dgmi_ctx* ctx, double* n, double* nlx, double* weight,
                        double* sn, double* snlx, double* sweight, int64_t* gi_align);

/* r = b - A x  (b may be NULL: r = -A x).  Replaces get_residual_or_smooth with
 * do_smooth=False (z02/volume_mf_diff.py:403-483; kappa as
 * z05/volume_mf_um_diff_varprop.py:347-436, 509-724). */
int dgmi_apply(dgmi_ctx* ctx, int dtype, const void* x, const void* b, void* r, void* stream);

/* r = b - A x of the FP64 operator (FP64 x, b, r; b required), every operation carried in
 * double-double arithmetic (about 106 bits) and the result rounded once to FP64.  The
 * accurate counterpart of dgmi_apply for the residuals that decide convergence: the FP64
 * evaluation rounds at about eps |A||x|, which for strongly varying diffusivity on fine
 * meshes reaches 1e-10 ||b||.  No counterpart in the reference, which evaluates the
 * residual as get_residual_or_smooth does (z02/volume_mf_diff.py:403-483). */
int dgmi_residual_dd(dgmi_ctx* ctx, const double* x, const double* b, double* r, void* stream);

/* b = sum_e M_e (f_e + alpha_e / dt) + Dirichlet-face terms with data u_bc, scaled by the
 * owner's kappa, + Neumann-face terms int v g_N -- _get_rhs (z02/volume_mf_diff.py:153-400:
 * volume :195-239 with the transient term :230-238, Dirichlet faces :298-357, Neumann faces
 * (glb_bcface_type 1) :360-400, not scaled by kappa there; kappa on Dirichlet faces as
 * z05/volume_mf_um_diff_varprop.py:320-342).  f, u_bc, b: FP64 device vectors in DOF layout;
 * u_neumann (nodal Neumann flux, the reference's u_bc[1]) and alpha_x_i (nodal data of the
 * transient term, used only when not NULL, with dt > 0) may be NULL. */
int dgmi_rhs(dgmi_ctx* ctx, const double* f, const double* u_bc, const double* u_neumann,
             const double* alpha_x_i, double dt, double* b, void* stream);

/* out[0] = l2, out[1] = linf (host doubles) of x_num - x_ana with the reference's
 * definition -- get_l2_error (z02/volume_mf_diff.py:2032-2056):
 *   l2^2 = sum_e sum_n sum_g (n[n,g] (x_num - x_ana)[e,n])^2 detwei[e,g],
 *   linf = max |x_num - x_ana|.
 * dtype as dgmi_apply; synchronizes the stream. */
int dgmi_l2_error(dgmi_ctx* ctx, int dtype, const void* x_num, const void* x_ana, double* out,
                  void* stream);

/* bdiagK (nele, nloc, nloc) and diagK (nele, nloc) in FP64 -- get_bdiag_diag
 * (z02/get_bdiag_diag.py:30-62).  Either output may be NULL. */
int dgmi_bdiag(dgmi_ctx* ctx, double* bdiag, double* diag, void* stream);

/* Build the per-element smoother matrices S (c = S r) for `variant`:
 * DGMI_SMOOTH_JACOBI3: three damped point-Jacobi sweeps on bdiagK from c = 0
 *   (z02/volume_mf_diff.py:474-480, blk_solver='jacobian');
 * DGMI_SMOOTH_DIRECT: S = bdiagK^-1 (z05/volume_mf_um_diff_varprop.py:384-392). */
int dgmi_smoother_setup(dgmi_ctx* ctx, int variant, double omega, void* stream);

/* One smoothing step: r = b - A x; x += omega * S r (get_residual_or_smooth with
 * do_smooth=True).  r_out may be NULL (the residual is then not stored). */
int dgmi_smooth(dgmi_ctx* ctx, int dtype, void* x, const void* b, void* r_out, void* stream);
/* n smoothing steps in a row (n >= 1), as n calls of dgmi_smooth without r_out (the
 * reference's smoothing loops, z02/volume_mf_diff.py:1947-1951).  The FP32 neighbour-form
 * sweeps of the vector kernels carry the face traces from one step to the next. */
int dgmi_smooth_n(dgmi_ctx* ctx, int dtype, void* x, const void* b, int n, void* stream);

/* Multigrid set-up (z02/volume_mf_diff.py:1642-1754, z02/multigrid_linearelastic.py:175-283):
 * P1CG sparsity, RAR (assembled from the continuous P1 element matrices by default; with
 * "rar_probing" 1, distance-2 colouring and coloured probing with the FP64 apply as the
 * reference), SFC ordering, SFC level operators.  Synchronises the stream. */
int dgmi_mg_setup(dgmi_ctx* ctx, void* stream);
/* info[0] = cg_nonods, info[1] = ncola, info[2] = ncolor (0 when RAR was assembled),
 * info[3] = nlevel,
 * info[4..4+nlevel) = nodes per level */
int dgmi_mg_info(const dgmi_ctx* ctx, int64_t* info, int cap);
/* host copies: fina (cg+1), cola (ncola), rar values (ncola), sfc (cg, 1-based) */
int dgmi_mg_export(const dgmi_ctx* ctx, int64_t* fina, int64_t* cola, double* rar,
                   int64_t* sfc);

/* P3DG -> P1CG restriction in SFC order: r1_sfc = perm(I_cf P^T r)
 * (z02/volume_mf_diff.py:1959-1996). */
int dgmi_restrict(dgmi_ctx* ctx, int dtype, const void* r_dg, void* r_cg_sfc, void* stream);
/* x += P I_fc e, with e in SFC order (z02/volume_mf_diff.py:2007-2018). */
int dgmi_prolong_add(dgmi_ctx* ctx, int dtype, const void* e_cg_sfc, void* x_dg, void* stream);
/* One SFC V-cycle on the P1CG levels, r and e in SFC order -- mg_on_P1CG
 * (z02/multigrid_linearelastic.py:112-172). */
int dgmi_sfc_vcycle(dgmi_ctx* ctx, int dtype, const void* r_sfc, void* e_sfc, void* stream);
/* One V-cycle from x = 0: x = M^-1 b -- _um_left_precond (z02/volume_mf_diff.py:1934-2029). */
int dgmi_precond(dgmi_ctx* ctx, int dtype, const void* b, void* x, void* stream);

/* Left-preconditioned restarted GMRES(m) for A x = b on device vectors (FP64 x, b)
 * -- _gmres_mg_solver (z02/volume_mf_diff.py:1757-1841).  inner_dtype selects the
 * arithmetic of the operator applies and V-cycles inside the Krylov loop; with
 * DGMI_F32 the solve runs as FP64 iterative refinement around FP32 cycles (flexible
 * GMRES unless "inner_flexible" is 0), with DGMI_BF16 likewise with bfloat16 inside the
 * preconditioner: its applies take bf16 inputs on the matrix cores, or, with
 * DGMI_SMOOTH_DIRECT and the parameter "bf16_smoother", its sweeps read a bfloat16 copy of
 * the smoother (FP32 arithmetic and face traces, vector kernels).  The Krylov operator
 * stays FP32.  Stops when ||b - A x|| <= rtol ||b|| (FP64 true residual)
 * or after max_its inner iterations, per component.  Synchronises the stream.  stats
 * (may be NULL), 4 values per component: iterations, final ||r||/||b||, cycles (outer
 * refinement steps), cycles run in FP64 arithmetic.  The true residual is evaluated as
 * "residual_dd" selects; FP32/BF16 refinement finishes in FP64 cycles as "fp64_tail"
 * selects (dgmi_set_param). */
int dgmi_gmres_solve(dgmi_ctx* ctx, int inner_dtype, const double* b, double* x, double rtol,
                     int m, int max_its, double* stats, void* stream);

/* MG-preconditioned conjugate gradients for A x = b on device vectors (FP64 x, b) --
 * _cg_mg_solver (z02/volume_mf_diff.py:1875-1931): r = b - A x, z = M^-1 r (one V-cycle,
 * as dgmi_precond), p = z; per iteration alpha = (r.z)/(p.Ap), x += alpha p,
 * r -= alpha Ap, z = M^-1 r, beta = (r.z)_new/(r.z)_old, p = z + beta p.  Stops, per
 * component, when the updated residual satisfies ||r|| <= abs_tol (abs_tol >= 0: the
 * reference's absolute test) or ||r|| <= rtol ||b|| (abs_tol < 0), or after max_its
 * iterations.  Synchronises the stream.  stats (may be NULL), 4 values per component:
 * iterations, final true ||b - A x|| / ||b||, 1, 1. */
int dgmi_cg_solve(dgmi_ctx* ctx, const double* b, double* x, double rtol, double abs_tol,
                  int max_its, double* stats, void* stream);

/* Level operations on caller-owned SFC levels (CSR with int32 rowptr/col, values and
 * diagonal of the level's dtype), for a caller that drives the V-cycle itself:
 * dgmi_level_smooth    rr = b - A e; e += omega rr / diag, with rr returned
 *                      (mg_smooth_one_level, z02/multigrid_linearelastic.py:25-39);
 * dgmi_level_restrict  coarse[I] = fine[2I] + fine[2I+1], fine padded with one zero
 *                      ((n_fine - 1)/2 + 1 coarse values; the restriction of r and e in
 *                      mg_on_P1CG, :147-153);
 * dgmi_level_prolong_add  fine[i] += coarse[i >> 1] (nearest upsampling, :163-164). */
int dgmi_level_smooth(int dtype, int64_t n, const int32_t* rowptr, const int32_t* col,
                      const void* val, const void* diag, double omega, void* e, const void* b,
                      void* rr, void* stream);
int dgmi_level_restrict(int dtype, int64_t n_fine, const void* fine, void* coarse, void* stream);
int dgmi_level_prolong_add(int dtype, int64_t n_fine, const void* coarse, void* fine,
                           void* stream);
/* Element-local P3DG <-> P1DG transfers with the P1 interpolation table at the P3 nodes:
 * direction 0 y[e][k] = sum_n P[n][k] x[e][n] (vel_pndg_to_p1dg_restrictor,
 * z02/multigrid_linearelastic.py:286-290), direction 1 y[e][n] = sum_k P[n][k] x[e][k]
 * (vel_p1dg_to_pndg_prolongator, :293-297); P from z02/mesh_init.py:224-235, 610-631. */
int dgmi_pn_p1dg(int D, int dtype, int direction, int64_t nele, const void* x, void* y,
                 void* stream);

/* Autotuner of the FP32 apply: times (reps launches each, on this context's mesh) the
 * vector kernels and the matrix-core kernels with 2, 4 and 8 waves (32, 64, 128 elements)
 * per work-group, fp32 and bf16-input, and keeps the fastest for this context.  Times the
 * smoothing step once the smoother is set up, the plain apply before.
 * choice[3] = f32 kernel family (0 vector, 1 MFMA), fp32 and bf16 waves per work-group;
 * times_us[7] = vector, fp32 x {2,4,8}, bf16 x {2,4,8}.  No counterpart in the reference. */
int dgmi_autotune(dgmi_ctx* ctx, int reps, int32_t* choice, double* times_us, void* stream);

/* Solver parameters by name:
 *   "gmres_orth"      0 = modified Gram-Schmidt, as the reference
 *                     (z02/volume_mf_diff.py:1813-1816); 1 = classical Gram-Schmidt with
 *                     one re-orthogonalisation (default).
 *   "residual_dd"     true residuals of dgmi_gmres_solve and the final one of
 *                     dgmi_cg_solve: 0 = FP64 (dgmi_apply); 1 = FP64 until an evaluation
 *                     would declare a component converged or shows that the last cycle
 *                     reduced its residual by less than half, then that evaluation and the
 *                     later ones in double-double (dgmi_residual_dd) (default); 2 = always
 *                     double-double.
 *   "fp64_tail"       FP32/BF16 refinement: 1 = once a cycle reduces a component's true
 *                     residual by less than half, the remaining cycles of the solve run in
 *                     FP64 (left-preconditioned, with the basis that fits in the FP32
 *                     workspace: 14 vectors for m = 30) (default); 0 = every cycle in the
 *                     inner precision.
 *   "inner_flexible"  FP32/BF16 refinement cycles: 1 = flexible (right-preconditioned)
 *                     GMRES (default), 0 = left-preconditioned like the FP64 solve.
 *   "f32_kernel"      kernels of FP32 applies: 0 = vector FMA (default), 1 = MFMA fp32,
 *                     2 = MFMA with bf16 inputs.
 *   "graph_vcycle"    1 = replay the V-cycle from a captured HIP graph, 0 = launch its
 *                     kernels one by one on the stream (default: measured 1% faster in a
 *                     solve, where the switch between a graph and stream work idles the
 *                     device for ~15 us per cycle).
 *   "smooth_offdiag"  FP32 sweeps of the exact block smoother (DGMI_SMOOTH_DIRECT) inside
 *                     the preconditioner: 1 = x = (1 - omega) x + S (b - A_off x), which
 *                     needs no volume term and only the neighbour half of the face terms
 *                     (default); 0 = x += S (b - A x).  Equal in exact arithmetic; FP64
 *                     sweeps and dgmi_smooth with r_out always use the second form.
 *   "bf16_smoother"   DGMI_BF16 inner cycles with DGMI_SMOOTH_DIRECT: 1 = the sweeps read a
 *                     bfloat16 copy of the smoother (half its bytes; iteration counts kept
 *                     at moderate diffusivity contrast, raised at high contrast); 0 = the
 *                     bf16-input matrix-core applies (default).
 *   "coarse_exact"    FP32 V-cycles: 1 = stop at one SFC level and solve it with its
 *                     dense inverse formed at dgmi_mg_setup (default): a context of one
 *                     solve uses the first level of at most 4096 rows, a packed context
 *                     the first level of at most 512 rows at or below each solve's LDS
 *                     start; 0 = sweep every level as the FP64 cycle (the reference's)
 *                     does.  dgmi_mg_info writes the level to info[4 + nlevel] when the
 *                     buffer holds it (-1: none).
 *   "f32_smooth_its"  pre- and post-smoothing sweeps of the FP32 cycles with carried traces
 *                     (default 3, the reference's pre_smooth_its, z02/config.py:112; 2 or 4
 *                     measured no faster overall).
 *   "dense_max_n"     FP32 cycles of a context of one solve: the exact level is the first SFC
 *                     level of at most this many rows (default 4096; set before
 *                     dgmi_mg_setup; 8192 and 16384 measured no fewer iterations).
 *   "single_pass"     vector kernels that form the face traces from the rows of x, own
 *                     and neighbours' (the neighbour's row relabelled into the face's own
 *                     frame), instead of a trace pass before them: 0 never (traces through
 *                     trace buffers), 1 the applies (dgmi_apply, the solvers' operator and
 *                     residuals), 2 also the FP32 preconditioner's sweeps (default).
 *   "smoother_half"   FP32 block-staged sweeps read an FP16 copy of the packed smoother
 *                     (each element's entries divided by its largest magnitude, kept in
 *                     FP32): 1 on, 0 off; by default on in 3D and off in 2D.  The copy is
 *                     formed by dgmi_smoother_setup when the parameter is on then (set it
 *                     before to use it in 2D); FP64 refinement keeps the final accuracy,
 *                     the iteration count can rise at high contrast.
 *   "rar_probing"     how dgmi_mg_setup forms RAR: 0 = assembled (default): on continuous
 *                     P1 functions the interior-face terms of the form vanish, so RAR[i][j]
 *                     sums the P1-reduced volume and Dirichlet-face element terms over the
 *                     elements holding both vertices, over mult_i mult_j (no colouring, no
 *                     operator applies); 1 = coloured probing with the FP64 apply as the
 *                     reference (z02/volume_mf_diff.py:1696-1754).  Equal to rounding.
 *   "staged"          FP32 single-pass applies, sweeps and the sweep from zero: 1 = the
 *                     block-staged kernels (rows read and written through LDS, face terms
 *                     in node form, volume term through the P2 gradient; default); 0 = the
 *                     per-wave single-pass kernels (same operator, other rounding).
 *   "destroy_sync"    dgmi_ctx_destroy: 1 = synchronise the device first (default); 0 =
 *                     the caller has finished all work it queued for the context (the
 *                     serving layer, which synchronises its job's stream).
 *   "alloc_limit_bytes"  device bytes this context may hold (0 = no limit, the default):
 *                     a later allocation that would pass it fails, and the entry point
 *                     that needed it returns -2 ("device allocation failed") without
 *                     launching anything; the serving layer caps each job at its
 *                     admitted budget with it.
 *   "mfma_waves", "bf16_waves"  waves (16 elements each) per work-group of the fp32 / bf16
 *                     matrix-core applies: 2, 4 (default) or 8. */
int dgmi_set_param(dgmi_ctx* ctx, const char* name, double value);

/* Krylov vector primitives on device vectors of n FP64 or FP32 values. */
int dgmi_dot(int dtype, int64_t n, const void* x, const void* y, double* out_dev, void* stream);
int dgmi_axpy(int dtype, int64_t n, double a, const void* x, void* y, void* stream);
/* Host-to-device copy through the context's pinned staging buffers: the npieces host
 * pieces src[i] (bytes[i] each) land one after the other from dst, queued on `stream`
 * (copy-engine DMA from pinned memory, which no kernel queue holds up).  Returns once
 * every byte has reached the device, so the host pieces may be reused at once.  No
 * counterpart in the reference, whose tensors torch creates on the device; the host
 * wrapper uploads right-hand-side data with it (per-request pieces of a packed job go
 * straight to their offsets, without a host concatenation). */
int dgmi_upload(dgmi_ctx* ctx, void* dst, int npieces, const void* const* src,
                const int64_t* bytes, void* stream);
/* Allocation-free dgmi_dot: out_dev[0] = x . y (FP64 accumulation) with the partial sums
 * in a fixed array of this library in device memory (256 KiB per GPU, held until the
 * process ends) instead of a buffer from hipMalloc.  No counterpart in the reference; it is
 * for callers that cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation.  n must lie in 0..dgmi_dot_noalloc_max_n() (2^28
 * values: 32768 partial sums of 8192 values each); larger n returns an error.  The first
 * call on a device looks up that device's copy of the array (no allocation).  Every call
 * on one GPU shares its array: calls that overlap on one GPU (other streams or host
 * threads) race.  The result is written before the stream's next work. */
int dgmi_dot_noalloc(int dtype, int64_t n, const void* x, const void* y, double* out_dev,
                     void* stream);
int64_t dgmi_dot_noalloc_max_n(void);

/* ---- host-only set-up (no device needed) ------------------------------------------
 * Replacements of the reference's f2py Fortran modules and Python set-up loops; also
 * used internally by dgmi_ctx_create / dgmi_mg_setup. */
/* element tables as dgmi_element_tables, without a context (SHATRInew,
 * z02/shape_function.py:459-1317) */
int dgmi_host_element_tables(int D, double* n, double* nlx, double* weight, double* sn,
                             double* snlx, double* sweight, int64_t* gi_align);
/* nbf / alnmt / glb_bcface_type (z02/mesh_init.py:293-390 with get_nb.getfinele,
 * z00-fortran-src/get_neighbouring.f90:1-130) */
int dgmi_host_topology(const int64_t* cells, int64_t nele, int D, int64_t* nbf, int64_t* alnmt,
                       int64_t* bctype);
/* P1CG CSR pattern (get_nb.getfin_p1cg, get_neighbouring.f90:133-156; z02/sparsity.py:102-122) */
int dgmi_host_p1cg_sparsity(const int64_t* cells, int64_t nele, int D, int64_t nverts,
                            int64_t cap, int64_t* fina, int64_t* cola, int64_t* ncola);
/* greedy distance-2 colouring, colours 1..ncolor (_color2, z02/sparsity.py:235-269) */
int dgmi_host_color2(const int64_t* fina, const int64_t* cola, int64_t n, int32_t* whichc,
                     int32_t* ncolor);
/* SFC numbering sfc[node] = 1-based curve position; replaces
 * sfc.ncurve_python_subdomain_space_filling_curve (combined_fortran_source.f90:10575)
 * with a Hilbert order of the vertex coordinates */
int dgmi_host_sfc_order(const double* verts, int64_t nverts, int D, int64_t* sfc);
/* calculate_nlevel_sfc (z00-fortran-src/compact_fcode.f90:229-250) */
int dgmi_host_nlevel_sfc(int64_t n);
/* SFC level operators, concatenated per level (best_sfc_mapping_to_sfc_matrix_unstructured,
 * compact_fcode.f90:7-225): level_n[l], rowptr_all (sum(level_n+1)), col_all/val_all */
int dgmi_host_sfc_levels(const int64_t* fina, const int64_t* cola, const double* a,
                         const int64_t* sfc, int64_t n, int64_t cap_rows, int64_t cap_nnz,
                         int64_t* nlevel, int64_t* level_n, int64_t* rowptr_all,
                         int64_t* col_all, double* val_all);

const char* dgmi_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DGMI_H */

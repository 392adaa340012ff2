/*
 * vv1d.h -- C ABI of libvv1d.so, the MI355X (gfx950) implicit timestep of the
 * 1D1V Vlasov-Poisson variational integrator (split mode, matrix-free,
 * nonlinear implicit midpoint; Arakawa J1/J2/J4 brackets; kinetic
 * FFT(x) x tridiagonal(v) preconditioner; GMRES; periodic Poisson).
 *
 * Conventions
 *  - Every field argument is a DEVICE pointer to float64 in the reference's
 *    PETSc natural order: element (i, j) (x index i, v index j) lives at
 *    i + nx*j, i.e. a C-contiguous (nv, nx) array.  A petsc4py Vec array of
 *    the reference's da1 maps onto it without a copy.
 *  - 1D x-arrays have nx entries, 1D v-arrays nv entries.
 *  - All work is enqueued on the context's stream; functions that return a
 *    host scalar synchronise that stream.
 *  - Every function returns 0 on success or a negative VV_ERR_* code;
 *    vv_last_error(ctx) (or vv_last_error(NULL) after a failed create)
 *    returns a message.
 *  - One context serves one stream; calls on one context must be serialised
 *    by the caller.  Separate contexts are independent.
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to the viVlasov1D source tree).
 */
#ifndef VV1D_H
#define VV1D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vv_ctx vv_ctx;

/* error codes */
#define VV_OK               0
#define VV_ERR_INVALID     -1   /* bad argument / unsupported size          */
#define VV_ERR_HIP         -2   /* HIP runtime error (launch, alloc, copy)  */
#define VV_ERR_UNSUPPORTED -3   /* option the library does not provide       */
#define VV_ERR_NONFINITE   -4   /* non-finite value met where one is fatal   */

/* Poisson bracket: PoissonBracket.create('arakawaj1'|'arakawaj2'|'arakawaj4')
 * (vlasov/solvers/components/PoissonBracket.pyx:37-50) */
#define VV_BRACKET_J1 1
#define VV_BRACKET_J2 2
#define VV_BRACKET_J4 4

/* preconditioner: TensorProductPreconditioner.create(None|'tensor_kinetic')
 * (vlasov/solvers/preconditioner/TensorProduct.pyx:66-86) */
#define VV_PRECOND_NONE           0
#define VV_PRECOND_TENSOR_KINETIC 1
#define VV_PRECOND_TENSOR_POTENTIAL 2   /* FFT along v x tridiagonal along x */

/* Laplacian: poisson module PETScPoissonCFD2 | PETScPoissonCFD4
 * (vlasov/run/run_base.py:93-94, 115) */
#define VV_LAPLACE_CFD2 2
#define VV_LAPLACE_CFD4 4

/* KSP converged reasons (PETSc numbering, as KSPGetConvergedReason) */
#define VV_KSP_CONVERGED_RTOL      2
#define VV_KSP_CONVERGED_ATOL      3
#define VV_KSP_DIVERGED_ITS       -3
#define VV_KSP_DIVERGED_DTOL      -4
#define VV_KSP_DIVERGED_BREAKDOWN -5
#define VV_KSP_DIVERGED_NANORINF  -9

/* Grid and operator selection.  Mirrors Grid.create
 * (vlasov/core/Grid.pyx:14-51) and the solver options of run_base.py:35-196. */
typedef struct {
    int32_t nx;            /* points in x (periodic); 8..4096 (any length:    */
                           /* FFT for powers of two, else a direct DFT)      */
    int32_t nv;            /* points in v; 8..4096                            */
    double  hx;            /* L / nx                                          */
    double  hv;            /* (vmax - vmin) / (nv - 1)                        */
    double  ht;            /* time step                                       */
    double  charge;        /* particle charge q (Poisson right-hand side)     */
    int32_t bracket;       /* VV_BRACKET_*                                    */
    int32_t precond;       /* VV_PRECOND_*                                    */
    int32_t laplace;       /* VV_LAPLACE_*                                    */
    int32_t restart;       /* GMRES restart length (PETSc default 30), <= 64  */
    const double *v;       /* HOST, nv: velocity grid v_j                     */
    const double *eigen;   /* HOST, re/im interleaved: the eigenvalues of
                              TensorProduct.compute_eigenvalues
                              (TensorProduct.pyx:109-116) -- of length nx for
                              VV_PRECOND_TENSOR_KINETIC, of length nv for
                              VV_PRECOND_TENSOR_POTENTIAL                     */
} vv_params;

/* Solver construction: PETScVlasovSolverBase.__init__
 * (vlasov/solvers/vlasov/PETScVlasovSolver.pyx:27-104) together with the
 * factored preconditioner (TensorProductKineticFast.pyx:25-95) and the
 * Poisson solver (PETScPoissonSolver.pyx:23-45).  stream: a hipStream_t, or
 * NULL for the null stream. */
int vv_ctx_create(const vv_params *p, void *stream, vv_ctx **out);

/* Batched context (no counterpart in the reference, which runs one instance
 * per process): `batch` independent instances that share nx, nv, the v-grid
 * (hv, v), dt, charge and the operator choice, but each with its own x-period
 * (hx[b], host array of `batch`; p->hx is ignored).  Every kernel runs all
 * instances in one launch (instance = second grid dimension).  Layout: a
 * field argument holds the instances back to back (instance b at b*nx*nv, a
 * C (batch, nv, nx) array); x-arrays likewise (b*nx); h0 and the v-arrays are
 * shared.  Functions that return host scalars return one per instance, in
 * instance order (vv_function_norm, vv_moments' pnorm: batch values;
 * vv_invariants: 6*batch; vv_ksp_solve / vv_snes_ksponly: batch results).
 * vv_norm2 / vv_dot / vv_axpy act on plain vectors of len.  The
 * mixed-precision solve takes one instance. */
int vv_ctx_create_batched(const vv_params *p, int32_t batch, const double *hx,
                          void *stream, vv_ctx **out);
int32_t vv_batch_size(const vv_ctx *ctx);

/* Instances that take part in the next linear solves (host array of `batch`
 * flags; NULL: all).  A masked instance's solve ends at once with reason 0 and
 * vv_snes_ksponly leaves its f unchanged (a finished Picard loop). */
int vv_set_active(vv_ctx *ctx, const int32_t *active);

/* Graph replay of vv_snes_ksponly: the residual, P^-1, GMRES set-up and the
 * first `restart` Arnoldi steps are captured once into a HIP graph (per f/fh
 * pair and tolerances) and replayed as one launch per Newton step; Arnoldi
 * steps past an instance's convergence return at once.  Needs a non-null
 * stream.  Restart cycles, if any instance needs one, run as ordinary
 * launches. */
int vv_set_graph(vv_ctx *ctx, int enable);

/* Launch-shape tuning (no counterpart in the reference): fft_rows = v-rows
 * per FFT workgroup (even, fft_rows * nx <= 8192), cgs_ept = 16-byte vectors
 * per thread of the Gram-Schmidt kernels (1, 2, 4, 8); 0 restores the
 * built-in choice.  vv_autotune times the candidates on the context's own
 * buffers (the FP64 preconditioner for fft_rows, mdot + maxpy at a
 * mid-cycle basis for cgs_ept), keeps the fastest and returns them; it
 * overwrites the context's work buffers and GMRES state. */
int vv_set_tuning(vv_ctx *ctx, int32_t fft_rows, int32_t cgs_ept);
int vv_get_tuning(vv_ctx *ctx, int32_t *fft_rows, int32_t *cgs_ept);
int vv_autotune(vv_ctx *ctx, int32_t *fft_rows, int32_t *cgs_ept);
/* Kernel fusion (no counterpart in the reference; on by default): where the
 * FFT workgroup's rows are exactly one Gram-Schmidt slice (power-of-two nx
 * <= 1024, kinetic preconditioner, rows per FFT workgroup x nx = 256 x
 * cgs_ept x 16 bytes / element size), the Arnoldi step runs the irFFT of
 * P^-1 J v and the CGS dots of the new vector in one launch, and where
 * besides nx >= 256 and no dissipation term is set, J v and its rFFT in one
 * launch (J v is not stored).  Results are identical either way (up to the
 * sign of exact zeros); 0 runs the separate kernels (for timing). */
int vv_set_fusion(vv_ctx *ctx, int enable);

/* Context lifetime and stream plumbing.  vv_ctx_destroy frees what the
 * context allocated, as PETScVlasovSolverBase.__dealloc__ destroys its local
 * Vecs (PETScVlasovSolver.pyx:105-111); never the caller's buffers.  The
 * others have no counterpart in the reference (PETSc keeps its own error
 * stack and runs on the host). */
int vv_ctx_destroy(vv_ctx *ctx);
const char *vv_last_error(const vv_ctx *ctx);
int vv_set_stream(vv_ctx *ctx, void *stream);
int vv_synchronize(vv_ctx *ctx);
/* bytes of device memory the context owns: what vv_ctx_create allocated
 * (Krylov basis, spectra, work vectors, ...) plus the buffers allocated on
 * first use (FP32 basis of the mixed solve, adjoint scratch, explicit-guess
 * stages, double-bracket fields) */
int64_t vv_scratch_bytes(const vv_ctx *ctx);

/* PETScVlasovSolverBase.update_previous (PETScVlasovSolver.pyx:142-157):
 * H_ave = ((((h0_j + .5 h1p_i) + .5 h1h_i) + .5 h2p_i) + .5 h2h_i), stored as
 * its 1D factors.  h0: nv entries; h1p, h1h, h2p, h2h: nx entries (device).
 * The context keeps a copy, as the reference keeps H_ave. */
int vv_set_hamiltonian(vv_ctx *ctx, const double *h0, const double *h1p,
                       const double *h1h, const double *h2p, const double *h2h);

/* PETScNLVlasovMP.function_solver (PETScNLVlasovMP.pyx:43-58):
 * y = (f - fh)/dt + B(.5 fh + .5 f, H_ave) on rows 2..nv-3, (f - fh)/dt on
 * the four v-boundary rows.  Overwrites y. */
int vv_residual(vv_ctx *ctx, const double *f, const double *fh, double *y);

/* PETScNLVlasovMP.jacobian_solver (PETScNLVlasovMP.pyx:33-40):
 * y = u/dt + .5 B(u, H_ave).  Overwrites y. */
int vv_jacobian_apply(vv_ctx *ctx, const double *u, double *y);

/* PoissonBracket.function (PoissonBracket.pyx:70-75, 325-330) with H = H_ave:
 * y += factor * {x, H_ave}.  Accumulates into y. */
int vv_bracket_add(vv_ctx *ctx, const double *x, double *y, double factor);

/* TensorProductPreconditionerKinetic.tensorProduct
 * (TensorProductKinetic.pyx:146-160, KineticFast.pyx:102-224):
 * y = irfft_x(T_k^-1 rfft_x(x)).  x and y may alias. */
int vv_precond_apply(vv_ctx *ctx, const double *x, double *y);

/* TensorProductPreconditioner.update_matrices, called from update_history at
 * every timestep (PETScVlasovSolver.pyx:135-139): for
 * VV_PRECOND_TENSOR_POTENTIAL the x-tridiagonal of every v-mode is rebuilt
 * from phi (an x-vector per instance; TensorProductPotentialFast.pyx:205-264):
 * sub e*1.0*(phi_i - phi_{i-1})*a, diagonal e*2.0*(phi_{i+1} - phi_{i-1})*a +
 * 1/dt, super e*1.0*(phi_{i+1} - phi_i)*a, a = 0.5/(12 hx hv), without the
 * periodic corners.  Until the first call the potential preconditioner is
 * dt * identity.  A no-op for VV_PRECOND_TENSOR_KINETIC, whose matrices do
 * not change. */
int vv_precond_update(vv_ctx *ctx, const double *phi);

/* PETScVlasovSolverBase.function / function_snes_mult / function_mult
 * (PETScVlasovSolver.pyx:168-173, 188-193): y = P^-1 F(f). */
int vv_function(vv_ctx *ctx, const double *f, const double *fh, double *y);

/* PETScVlasovSolverBase.jacobian / mult / snes_mult / jacobian_mult
 * (PETScVlasovSolver.pyx:160-165, 176-186): y = P^-1 J u. */
int vv_jacobian(vv_ctx *ctx, const double *u, double *y);

/* PETScVlasovSolverBase.function_norm (PETScVlasovSolver.pyx:196-198):
 * y = F(f) (unpreconditioned), *out = ||y||_2 (host). */
int vv_function_norm(vv_ctx *ctx, const double *f, const double *fh, double *y,
                     double *out);

/* run_base.calculate_residual (run_base.py:631-637) in one host read:
 * y = F(f), *fnorm = ||y||_2 and *pnorm = ||A phi + (n - mean n) q||_2 of
 * the last vv_moments call (per instance in a batched context). */
int vv_residual_norms(vv_ctx *ctx, const double *f, const double *fh, double *y,
                      double *fnorm, double *pnorm);

/* Toolbox.compute_density (vlasov/toolbox/Toolbox.pyx:79-98):
 * n_i = hv * sum_j f_ij. */
int vv_density(vv_ctx *ctx, const double *f, double *n);

/* Toolbox.compute_velocity_density / compute_energy_density
 * (Toolbox.pyx:103-148): u_i = hv sum_j v_j f_ij, e_i = hv sum_j v_j^2 f_ij
 * (x-vectors, per instance).  Output moments of the HDF5 schema; not on the
 * timestep's path. */
int vv_velocity_moments(vv_ctx *ctx, const double *f, double *u, double *e);

/* Dissipation term of the solver (PETScVlasovSolver.pyx:39-42, 101-102;
 * chosen by config.py:148-163: dissipation_type 'collisions' with
 * collision_operator collT / collE, or 'double_bracket' with bracket_operator
 * ffh / fhh; nothing when coll_freq == 0).  Once set, every residual
 * (vv_residual, vv_function, vv_function_norm, vv_snes_ksponly) adds the
 * term function_solver adds (PETScNLVlasovMP.pyx:56-58) and every Jacobian
 * apply (vv_jacobian_apply, vv_jacobian, the GMRES operator) the term
 * jacobian_solver adds (:38-39).  The double bracket's Jacobian uses the
 * average .5 fh + .5 f of the last residual evaluation, as the reference's
 * jacobian_solver uses self.Fave.  Single-instance FP64 contexts only: the
 * mixed-precision, adjoint and batched entry points return
 * VV_ERR_UNSUPPORTED while a dissipation term is set. */
#define VV_DISSIPATION_NONE  0
#define VV_DISSIPATION_COLLT 1   /* CollisionOperatorT (CollisionOperator.pyx:68-156) */
#define VV_DISSIPATION_COLLE 2   /* CollisionOperatorE (CollisionOperator.pyx:160-250) */
#define VV_DISSIPATION_FFH   3   /* DoubleBracketFFH (DoubleBracket.pyx:67-82) */
#define VV_DISSIPATION_FHH   4   /* DoubleBracketFHH (DoubleBracket.pyx:86-97) */

typedef struct {
    int32_t kind;          /* VV_DISSIPATION_*                                 */
    int32_t pad;
    double  coll_freq;     /* nu (solver.coll_freq)                            */
    double  coll_diff;     /* weight of the diffusion term (default 1)         */
    double  coll_drag;     /* weight of the drag term (default 1)              */
} vv_dissipation;

/* CollisionOperator.create / DoubleBracket.create as PETScVlasovSolverBase
 * calls them (PETScVlasovSolver.pyx:101-102): the term every later residual
 * and Jacobian apply adds (kind VV_DISSIPATION_NONE: none). */
int vv_set_dissipation(vv_ctx *ctx, const vv_dissipation *d);

/* PETScVlasovSolverBase.set_moments (PETScVlasovSolver.pyx:123-132): device
 * x-vectors (nx each) of the current (p) and history (h) density N, velocity
 * density U, energy density E and collision factor A.  The context keeps the
 * pointers, not copies, as the reference keeps the Vec objects: the caller
 * refreshes their contents (vv_velocity_moments, vv_collision_factor) and
 * keeps them alive.  E may be NULL (no operator reads it). */
int vv_set_moments(vv_ctx *ctx, const double *np, const double *up, const double *ep,
                   const double *ap, const double *nh, const double *uh,
                   const double *eh, const double *ah);

/* Toolbox.compute_collision_factor (Toolbox.pyx:153-165):
 * a_i = n_i n_i / (n_i e_i - u_i u_i). */
int vv_collision_factor(vv_ctx *ctx, const double *n, const double *u, const double *e,
                        double *a);

/* CollisionOperator{T,E}.function (CollisionOperator.pyx:72-110, 164-202) of
 * the context's collision operator: y += C(f) on rows 2..nv-3 with
 * u = U/n.  e is unused (signature parity) and may be NULL. */
int vv_collision_add(vv_ctx *ctx, const double *f, double *y, const double *n,
                     const double *u, const double *e, const double *a, double factor);

/* DoubleBracket{FFH,FHH}.function / .jacobian (DoubleBracket.pyx:69-97) of the
 * context's double bracket with H = H_ave: y += D(fave) resp. y += D'(fave) u. */
int vv_double_bracket_function(vv_ctx *ctx, const double *fave, double *y, double factor);
int vv_double_bracket_jacobian(vv_ctx *ctx, const double *u, const double *fave, double *y,
                               double factor);

/* Explicit Arakawa steppers of the initial guess (run_base.py:786-927,
 * chosen by solver.initial_guess, iterated solver.initial_iter = niter times
 * per step with dt/niter each).  The reference's PETScArakawaSymplectic and
 * PETScArakawaRungeKutta cannot run as written (bracket arguments in the
 * wrong order, an accumulator never cleared, an undefined rk18_J4); these
 * are the schemes they state, with the tendency of the implicit residual,
 * df/dt = -{f, H} on rows 2..nv-3.  x is updated in place; h0 is the v-vector
 * of the kinetic Hamiltonian, h1 and h2 x-vectors (the run script passes its
 * history potentials h1h, h2h).  Single-instance contexts. */

/* PETScArakawaSymplectic.kinetic (PETScArakawaSymplectic.pyx:62-77):
 * x += factor*ht/niter * (-{x, H0}). */
int vv_symplectic_kinetic(vv_ctx *ctx, double *x, const double *h0, double factor,
                          int32_t niter);

/* PETScArakawaSymplectic.potential (PETScArakawaSymplectic.pyx:80-97):
 * x += factor*ht/niter * (-{x, H1 + H2}). */
int vv_symplectic_potential(vv_ctx *ctx, double *x, const double *h1, const double *h2,
                            double factor, int32_t niter);

/* PETScArakawaRungeKutta.rk4_16 / rk4_18 (PETScArakawaRungeKutta.pyx:20-98):
 * one RK4 step (classical 1/6 rule or the 3/8 rule) of dt/niter with
 * H = H0 + H1 + H2. */
#define VV_RK4_16 1
#define VV_RK4_18 2
int vv_rk4(vv_ctx *ctx, double *x, const double *h0, const double *h1, const double *h2,
           int32_t rule, int32_t niter);

/* PETScPoissonSolver.formRHS + the Poisson KSPSolve
 * (PETScPoissonSolver.pyx:58-74; run_base.py:689-690): solves A phi = b,
 * b_i = -(n_i - mean n) q, exactly on the periodic grid; phi is returned
 * mean-free (the reference's null-space constant never reaches f). */
int vv_poisson_solve(vv_ctx *ctx, const double *n, double *phi);

/* PETScPoisson{CFD4,CFD2}.matrix_mult (PETScPoissonCFD4.pyx:64-78) */
int vv_poisson_matrix_mult(vv_ctx *ctx, const double *x, double *y);

/* PETScPoisson{CFD4,CFD2}.function_mult (PETScPoissonCFD4.pyx:83-100):
 * y = A x + (n - mean n) q. */
int vv_poisson_function_mult(vv_ctx *ctx, const double *x, const double *n,
                             double *y);

/* Toolbox.potential_to_hamiltonian (Toolbox.pyx:38-74), kept 1D:
 * h1_i = phi_i - sum(phi)/nx. */
int vv_potential_to_hamiltonian(vv_ctx *ctx, const double *phi, double *h1);

/* run_base.calculate_moments(potential=True) + copy_pint_to_h
 * (run_base.py:640-704, 741-742) fused: n from f, phi from n, h1 from phi,
 * and *pnorm = ||A phi + (n - mean n) q||_2 (host, may be NULL). */
int vv_moments(vv_ctx *ctx, const double *f, double *n, double *phi, double *h1,
               double *pnorm);

/* vv_moments without any allocation -- no counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same results as
 * vv_moments bit for bit (same kernels), but the density partial sums and the
 * Poisson work vectors live in a file-scope device array of the library
 * (2.125 MiB on every GPU the library runs on, held until the process ends)
 * instead of the context's buffers, and the residual norm
 * ||A phi + (n - mean n) q||_2 of instance b goes to pnorm[b], a DEVICE array
 * of batch doubles (may be NULL), so nothing is copied to the host.
 * Limit: nx * batch <= 4096; larger contexts get VV_ERR_INVALID.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (several
 * streams or host threads) race and must be serialised by the caller. */
int vv_moments_noalloc(vv_ctx *ctx, const double *f, double *n, double *phi, double *h1,
                       double *pnorm);

/* VecNorm(NORM_2), VecDot, VecAXPY over len doubles. */
int vv_norm2(vv_ctx *ctx, const double *x, int64_t len, double *out);
int vv_dot(vv_ctx *ctx, const double *x, const double *y, int64_t len, double *out);
int vv_axpy(vv_ctx *ctx, double alpha, const double *x, double *y, int64_t len);

/* PETSc KSPGMRES on the operator P^-1 J of this context (the MatPython
 * context set at run_nonlinear_matrixfree_split_mp.py:42-53): restart
 * `restart`, classical Gram-Schmidt, Givens rotations, x0 = 0, stop when
 * ||r_k|| <= max(rtol ||b||, atol) (preconditioned residual), or
 * ||r_k|| >= dtol ||b||, or max_it iterations. */
typedef struct {
    double  rtol;
    double  atol;
    double  dtol;
    int32_t max_it;
} vv_ksp_params;

typedef struct {
    int32_t its;       /* KSPGetIterationNumber        */
    int32_t reason;    /* KSPGetConvergedReason        */
    double  rnorm;     /* last residual estimate       */
    double  rnorm0;    /* ||b||                        */
} vv_ksp_result;

int vv_ksp_solve(vv_ctx *ctx, const double *b, double *x, const vv_ksp_params *kp,
                 vv_ksp_result *res);

/* SNESSolve_KSPONLY as set up at run_nonlinear_matrixfree_split_mp.py:47-53:
 * b = P^-1 F(f); KSPSolve(P^-1 J, b) -> delta; f -= delta unless the linear
 * solve diverged.  f is updated in place.  With VV_PRECOND_NONE P is the
 * identity (the reference's preconditioner_type None). */
int vv_snes_ksponly(vv_ctx *ctx, double *f, const double *fh,
                    const vv_ksp_params *kp, vv_ksp_result *res);

/* One timestep of every instance of a batched context, the whole Picard
 * loop in native code (no counterpart in the reference, which advances one
 * instance per process; the rules are its run loop's,
 * run_nonlinear_matrixfree_split_mp.py:111-174 and run_base.py:631-763,
 * applied per instance): make_history (fh = f, h1h = h1, h2h = h2,
 * phih = phi), update_previous, the initial residual ||F(f)|| + ||Poisson
 * residual||; then, while any instance is active, with the others masked
 * (vv_set_active): fl = f, vv_snes_ksponly, moments (density, phi, h1),
 * update_previous, the residual; an instance stops when its residual rose
 * (after the first iteration), fell below snes_atol or took snes_max_iter
 * iterations, and a rise restores its f from fl (moments and H again).
 * The caller's Python or C driver thus holds no lock between the launches.
 * history (optional, history_rows x batch doubles): the residual of every
 * iteration per instance, NaN where the instance was not active. */
typedef struct {
    double *f, *fh, *fl, *fy;    /* fields (batch, nv, nx): current, history, last iterate, work */
    const double *h0;            /* nv */
    double *h1, *h1h, *h2, *h2h; /* (batch, nx): factors of H, current and history */
    double *n, *phi, *phih;      /* (batch, nx): density, potential, its history */
} vv_batch_fields;

typedef struct {
    int32_t picard;    /* Picard iterations taken          */
    int32_t reverted;  /* the last iterate was undone      */
    int64_t gmres;     /* GMRES iterations over them       */
    double  initial;   /* residual before the loop         */
    double  final;     /* residual at its end              */
} vv_instance_step;

int vv_batched_timestep(vv_ctx *ctx, const vv_batch_fields *s, const vv_ksp_params *kp,
                        double snes_atol, int32_t snes_max_iter, vv_instance_step *rec,
                        double *history, int32_t history_rows, int32_t *iterations);

/* Mixed precision (no counterpart in the reference, whose solve is FP64
 * throughout; it replaces the same KSPSolve / SNESSolve_KSPONLY call sites):
 * iterative refinement on (P^-1 J) x = b.  Each correction equation is solved
 * by one FP32 GMRES cycle (FP32 Jacobian-apply, FP32 preconditioner, FP32
 * Krylov basis; dots, norms, Hessenberg and Givens in FP64) to a relative
 * reduction of inner_rtol (less when less meets the target); the residual
 * r = b - P^-1 J x, the solution x and the stopping test
 * ||r|| <= max(rtol ||b||, atol) stay FP64, on the true preconditioned
 * residual.  res->its counts FP32 iterations, res->rnorm is the true final
 * residual norm; *cycles (may be NULL) the FP32 cycles run.
 *
 * fp64_check = 0 is the form for an outer loop that evaluates the FP64
 * residual itself (the Picard iteration of the timestep, which then acts as
 * the refinement): the first FP32 cycle runs with the KSP's own rtol / atol
 * / dtol / max_it and, when it ends with a reason, that reason, its count
 * and its (FP64-accumulated) residual estimate are the result, with no FP64
 * operator applied.  Only a cycle that exhausts the restart length without
 * a reason continues with the FP64-checked refinement above. */
typedef struct {
    double  inner_rtol;    /* relative reduction per FP32 cycle, e.g. 1e-5  */
    int32_t max_cycles;    /* cap on FP32 cycles per solve                  */
    int32_t fp64_check;    /* 1: FP64 residual after every cycle; 0: see above */
} vv_mixed_params;

int vv_ksp_solve_mixed(vv_ctx *ctx, const double *b, double *x, const vv_ksp_params *kp,
                       const vv_mixed_params *mp, vv_ksp_result *res, int32_t *cycles);

/* vv_snes_ksponly with the linear solve done by vv_ksp_solve_mixed;
 * b = P^-1 F(f) and the update f -= delta are FP64. */
int vv_snes_ksponly_mixed(vv_ctx *ctx, double *f, const double *fh, const vv_ksp_params *kp,
                          const vv_mixed_params *mp, vv_ksp_result *res, int32_t *cycles);

/* The FP32 operators of the mixed solver on FP32 device fields (same layout):
 * y = u/dt + .5 B(u, H_ave) and y = P^-1 x.  x and y may alias in the
 * preconditioner. */
int vv_jacobian_apply32(vv_ctx *ctx, const float *u, float *y);
int vv_precond_apply32(vv_ctx *ctx, const float *x, float *y);

/* Linearisation and adjoint of the implicit timestep (no counterpart in the
 * reference, which has no derivatives).  With
 *   G(f1, f0) = (f1 - f0)/dt + R {(f0 + f1)/2, H},  H = h0 + h1[f1]/2 + h1[f0]/2,
 * R = keep rows 2..nv-3, h1[f] the h1 of vv_moments (linear in f) and
 * H_ave of vv_set_hamiltonian standing for H, about a converged step
 * (f0, f1) with fbar = (f0 + f1)/2:
 *   A = dG/df1 = I/dt + 1/2 R {., H} + 1/2 R {fbar, h1[.]},
 *   A0 = dG/df0 = -I/dt + (the same two terms).
 * The transposes use the skew-symmetry of the Arakawa stencils:
 *   A^T u = u/dt - 1/2 {R^T u, H} + 1/2 h1[{R^T u, fbar}] on all rows.
 * fbar = NULL drops the field-coupling term (A is then the Picard
 * Jacobian J).  vv_linear_apply: y = A u, A0 u, A^T u or A0^T u.
 * vv_precond_apply_t: y = P^-T x (mode blocks T_k^H).
 * vv_adjoint_solve: GMRES on P^-T A^T mu = P^-T lam (same KSP options).
 * vv_adjoint_step: mu from vv_adjoint_solve, then lam0 = -A0^T mu, the
 * sensitivity with respect to f0 of a function whose sensitivity with
 * respect to f1 is lam1. */
#define VV_LIN_A   0
#define VV_LIN_A0  1
#define VV_LIN_AT  2
#define VV_LIN_A0T 3
int vv_linear_apply(vv_ctx *ctx, const double *u, const double *fbar, double *y, int32_t which);
int vv_precond_apply_t(vv_ctx *ctx, const double *x, double *y);
int vv_adjoint_solve(vv_ctx *ctx, const double *lam, const double *fbar, double *mu,
                     const vv_ksp_params *kp, vv_ksp_result *res);
int vv_adjoint_step(vv_ctx *ctx, const double *lam1, const double *fbar, double *lam0,
                    double *mu, const vv_ksp_params *kp, vv_ksp_result *res);

/* Invariants of core/distribution_function.pyx:130-199,
 * core/hamiltonian.pyx:67-100 and core/potential.pyx:48-59:
 * out[0] mass, out[1] L2, out[2] total energy, out[3] kinetic energy,
 * out[4] potential energy, out[5] field energy (host array of 6). */
int vv_invariants(vv_ctx *ctx, const double *f, const double *h0,
                  const double *h1, const double *h2, const double *phi,
                  double *out);

/* Per-kernel timing (no counterpart in the reference) with HIP events attached to each timed kernel's own dispatch (for the
 * roofline report).  kernel: "jacobian", "residual", "rfft", "tridiag",
 * "irfft", "mdot", "maxpy_norm", the FP32 kernels of the mixed solver
 * "jacobian32", "rfft32", "tridiag32", "irfft32", "mdot32", "maxpy_norm32",
 * the fused irFFT + Gram-Schmidt dots "irfft_mdot", "irfft_mdot32", and
 * the Jacobian-vector product fused with the rFFT "jac_rfft", "jac_rfft32".
 * Enable, run, then read the summed time in ms, the launch count and the
 * summed algorithmic HBM bytes of those launches (Krylov launches that
 * returned at once after convergence are not counted).  Reading clears that
 * kernel's records.  enable is a bit mask over the kernels in the order
 * listed (jacobian = 1, residual = 2, rfft = 4, ..., maxpy_norm = 64,
 * jacobian32 = 128, ..., maxpy_norm32 = 4096, irfft_mdot = 8192,
 * irfft_mdot32 = 16384, jac_rfft = 32768, jac_rfft32 = 65536; -1 all, 0 off). */
int vv_profile_enable(vv_ctx *ctx, int enable);
/* Time one launch in every `every` (>= 1, default 1) of each enabled kernel
 * kind: a HIP-event-timed dispatch has idle time of its own around it, so a
 * sample keeps the measured run close to an untimed one. */
int vv_profile_sample(vv_ctx *ctx, int every);
int vv_profile_read(vv_ctx *ctx, const char *kernel, double *ms, int64_t *launches,
                    double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* VV1D_H */

// Internal declarations shared by the HIP translation units of libvv1d.so.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vv1d.h"
#include "tile_index.h"

namespace vv {

constexpr int kMaxRestart = 64;     // GMRES restart limit (basis of restart+1 vectors)
constexpr int kMaxNx = 4096;        // largest x size
constexpr int kMaxNv = 8192;
constexpr int kReduceBlocks = 512;  // fixed block count of N-sized reductions
constexpr int kDensityChunks = 64;  // row chunks of the density partial sums (DCH x nx per instance)

// Host-visible part of the GMRES state; copied to pinned host memory after
// each chunk of iterations.
struct GmresStatus {
    int32_t it;        // iterations in the current cycle
    int32_t its;       // total iterations
    int32_t reason;    // KSP converged reason (0 = iterating)
    int32_t done;      // current cycle finished: every Krylov kernel returns at once
    int32_t active;    // instance takes part in this solve (batched contexts)
    int32_t pad;
    double res;        // current residual estimate
    double rnorm0;     // norm of the right-hand side
};

// Per-instance grid constants (a batched context holds instances that share
// nx, nv, v-grid and dt but not the x-period, so hx differs).
struct InstConst {
    double hx, hx_inv, hx2_inv;
    double af1, af2;   // Arakawa factors hx^-1 hv^-1 / 12, / 24
    double afpc;       // preconditioner band factor 0.5 hx^-1 hv^-1 / 12
};

// Device-resident GMRES state.  Written only by the one-workgroup control
// kernels (cycle start, mdot finish, Hessenberg update, back substitution);
// read by every Krylov kernel.
struct GmresState {
    GmresStatus st;
    double hh[(kMaxRestart + 1) * kMaxRestart];  // rotated Hessenberg, column it at hh[it*(kMaxRestart+1)]
    double cc[kMaxRestart];
    double ss[kMaxRestart];
    double grs[kMaxRestart + 1];
    double scale[kMaxRestart + 1];   // v_j = scale[j] * V_j (V_j stored unnormalised)
    double coef[kMaxRestart + 1];    // coefficients of the pending maxpy
    double ttol, rtol, atol, dtol, haptol;
    int32_t max_it;
    int32_t max_k;
    int32_t first;                   // first cycle: sets rnorm0 and ttol
    int32_t pad;
};

}  // namespace vv

struct vv_ctx {
    vv_params p;
    hipStream_t stream;
    int B;              // instances (1 unless created batched); instance b of every
                        // per-instance buffer lives at base + b * stride
    vv::InstConst *inst;        // device, B entries
    std::vector<vv::InstConst> inst_host;
    int32_t *active;            // device, B flags: instance takes part in the next solve
    int64_t N;          // nx * nv (per instance)
    int64_t vs;         // element stride of the Krylov basis vectors: B N rounded up to a multiple of 4
    // transform geometry of the tensor-product preconditioner: the real FFT
    // has length fft_n over fft_rows rows and nmodes = fft_n/2 + 1 modes, each
    // mode a tridiagonal system of length fft_rows.  kinetic: FFT along x
    // (fft_n = nx, fft_rows = nv); potential: along v on the transposed field
    // (fft_n = nv, fft_rows = nx)
    int nmodes;         // fft_n/2 + 1
    int logn;           // log2(fft_n) when fft_n is a power of two (FFT kernels); else a direct DFT
    int fft_n, fft_rows;
    int64_t coef_stride;        // instance stride of gm/g0/gp (0: shared)
    double cf_lo, cf_mid, cf_hi;  // band constants: kinetic 0.5/1.0/0.5, potential 1.0/2.0/1.0
    double *tb;                 // potential: the transposed field (B N doubles)
    // grid-derived constants (x-dependent ones per instance in inst)
    double ht_inv, hv_inv;
    // Hamiltonian factors (device, ctx-owned copies): h0 shared, h1*/h2* per instance
    double *h0, *h1p, *h1h, *h2p, *h2h;
    // preconditioner data (device)
    double2 *eigen;     // nmodes used (mode k = 0..fft_n/2)
    double *gm, *g0, *gp;   // per-row real factors of the bands (kinetic: of v; potential: phi differences per instance)
    double *vgrid;          // v_j (velocity moments)
    double2 *twiddle;   // exp(-2 pi i m / fft_n), m = 0..fft_n-1
    float2 *twiddle32;  // the same, rounded to float (FP32 preconditioner)
    double2 *spec;      // nmodes * fft_rows complex spectrum per instance, mode-major
    // kinetic preconditioner: reciprocal LU pivots 1/u_j of every T_k, same
    // layout as the spectrum (FP64 at creation; FP32 on first FP32 use);
    // null for the potential preconditioner (its bands change every step)
    double2 *lu64;
    float2 *lu32;
    // Poisson
    double *pinv;       // first row of the mean-free circulant inverse (nx per instance)
    double *pwork;      // Poisson scratch (3*nx + 1 per instance)
    // work vectors (N)
    double *X;          // bracket / residual output before P^-1
    double *rhs;        // SNES right-hand side P^-1 F
    double *D;          // Newton update
    double *basis;      // (restart+1) vectors of B*N: V_j of instance b at basis + (j*B + b)*N
    // mixed precision (allocated on first use)
    float *basis32;     // FP32 basis, same layout
    double *R;          // FP64 true residual of the refinement loop
    double *w64, *d64;  // FP32 cycle with a dissipation term: the Krylov vector widened, the term in FP64
    // reductions
    double *partial;    // per instance: part_stride doubles
    int64_t part_stride;
    double *dscalar;    // device scalars, 16 per instance
    double *npart;      // density partial sums, 64 nx per instance
    vv::GmresState *gs;             // B states
    vv::GmresStatus *host_status;   // pinned, 2 slots of B statuses
    vv::GmresStatus *host_status_dev;   // the same memory as the device addresses it
    vv::GmresStatus *status_out;    // set for one launch: the Hessenberg update also stores the statuses there
    double *host_scalars;           // pinned staging of device scalars
    double *host_scalars_dev;       // the same memory as the device addresses it
    int32_t *host_active;           // pinned staging of the instance mask
    int32_t *host_active_dev;       // the same memory as the device addresses it
    hipEvent_t ev[2];
    // profiling (per-kernel HIP event pairs)
    int prof;           // bit k set: time kernel kind k
    void *prof_state;
    hipEvent_t prof_a, prof_b;   // events of the next main-kernel launch (null: none)
    // linearised / adjoint operators (scratch allocated on first use):
    // operator of the GMRES cycle (0: P^-1 J of the Picard step; 1: P^-T A^T,
    // the transposed full linearisation), the midpoint state f_bar of its
    // field coupling (null: no coupling)
    int op_mode;
    const double *op_fbar;
    double *T2;         // B*N
    double *xs;         // 4 regions of B*nx: n | phi | h1 | spare
    // launch-shape tuning (0: built-in heuristic): v-rows per FFT workgroup,
    // 16-byte vectors per thread of the CGS kernels
    int tune_fft_rows, tune_cgs_ept;
    // kernel fusion (default on): the Arnoldi step's irFFT and CGS dots in
    // one launch where the geometry allows; results are bit-identical
    bool fuse;
    // captured first GMRES cycle of SNESSolve_KSPONLY (graph replay mode)
    bool use_graph, capturing;
    hipGraphExec_t graph_exec;
    const double *graph_f, *graph_fh;
    double graph_kp[4];
    // dissipation term (VV_DISSIPATION_*) and its parameters; moment vectors
    // are caller-owned device x-vectors: N, U, E, A current then history
    int diss_kind;
    double coll_freq, coll_diff, coll_drag, hv2_inv;
    const double *mom[8];
    double *fave;       // .5 fh + .5 f of the last residual (double brackets)
    double *db1;        // {fave, H_ave}, valid while db1_valid
    double *db2;        // scratch bracket of the Jacobian
    bool db1_valid;
    double *ex;         // explicit steppers: H | t1 | t2 | t3 | t4 | stage (6 N)
    int64_t dev_bytes;  // device memory the context has allocated so far
    double *static_scratch;   // device address of this device's copy of the allocation-free
                              // entry point's file-scope scratch array
    std::string err;
};

namespace vv {

// ---- launchers -------------------------------------------------------------
// bracket.hip
hipError_t launch_residual(vv_ctx *c, const double *f, const double *fh, double *y);
hipError_t launch_jacobian(vv_ctx *c, const double *u, const double *uscale, double *y);
hipError_t launch_bracket_add(vv_ctx *c, const double *x, double *y, double factor);
hipError_t launch_jacobian32(vv_ctx *c, const float *u, const double *uscale, float *y);
hipError_t launch_widen(vv_ctx *c, const float *x, double *y);                  // y = (double) x
hipError_t launch_narrow_add(vv_ctx *c, float *y, const double *d);           // y = (float)(y + d)
// y = alpha u + factor {u~, H} on all rows (u~: u on rows 2..nv-3, else 0);
// H = h2d (2D) or hxv (x only) when given, else H_ave
hipError_t launch_bracket_t(vv_ctx *c, const double *u, const double *uscale, const double *h2d, const double *hxv,
                            double *y, double alpha, double factor);
// y += factor {x, H} on rows 2..nv-3, H = h2d / hxv / H_ave as above
hipError_t launch_bracket_gen(vv_ctx *c, const double *x, const double *h2d, const double *hxv, double *y,
                              double factor);
// y (+)= factor {s x, H} on rows 2..nv-3 (store: y = 0 + ..., other rows 0),
// s = *xscale when given, H = h2d or H_ave
hipError_t launch_bracket_acc(vv_ctx *c, const double *x, const double *xscale, const double *h2d, double *y,
                              double factor, bool store, const int32_t *done);
// dissipation.hip: y += C(s f) of the context's collision operator; a = n n / (n e - u u)
hipError_t launch_collision(vv_ctx *c, const double *f, const double *fscale, double *y, const double *n,
                            const double *U, const double *a, double factor, const int32_t *done);
hipError_t launch_collision_factor(vv_ctx *c, const double *n, const double *u, const double *e, double *a);
hipError_t launch_average(vv_ctx *c, const double *f, const double *fh, double *out);   // .5 fh + .5 f
// explicit.hip: H = h0 | h1 + h2 | (h0 + h1) + h2 as a 2D field; stage combinations
hipError_t launch_compose_h(vv_ctx *c, const double *h0, const double *h1, const double *h2, double *H);
hipError_t launch_stage(vv_ctx *c, const double *x, const double *t, double s, double *out);   // x + s t
hipError_t launch_rk18_stage3(vv_ctx *c, const double *x, const double *t1, const double *t2, double cc, double s,
                              double *out);
hipError_t launch_rk18_stage4(vv_ctx *c, const double *x, const double *t1, const double *t2, const double *t3,
                              double s, double *out);
hipError_t launch_rk4_end(vv_ctx *c, bool three_eighths, double *x, const double *t1, const double *t2,
                          const double *t3, const double *t4, double s);
// spectral.hip
int fft_blocks(vv_ctx *c);
hipError_t launch_rfft(vv_ctx *c, const double *x, double2 *spec, const int32_t *done);
// J u (u times its optional device scale) and its rFFT in one launch, the
// Jacobian never stored; jac_rfft_rows: the rows per workgroup where the
// fused kernel applies to the context (the kinetic form with a power-of-two
// nx in [256, 1024]), else 0.  No dissipation terms (the caller checks).
int jac_rfft_rows(const vv_ctx *c);
hipError_t launch_jac_rfft(vv_ctx *c, const double *u, const double *uscale, double2 *spec, const int32_t *done);
hipError_t launch_jac_rfft32(vv_ctx *c, const float *u, const double *uscale, float2 *spec, const int32_t *done);
hipError_t launch_tridiag_k(vv_ctx *c, double2 *spec, const int32_t *done);
hipError_t launch_irfft(vv_ctx *c, const double2 *spec, double *y, const int32_t *done);
hipError_t launch_tridiag_t(vv_ctx *c, double2 *spec, const int32_t *done);   // T_k^-H (transposed P)
hipError_t launch_rfft32(vv_ctx *c, const float *x, float2 *spec, const int32_t *done);
hipError_t launch_tridiag32(vv_ctx *c, float2 *spec, const int32_t *done);
hipError_t launch_irfft32(vv_ctx *c, const float2 *spec, float *y, const int32_t *done);
// irFFT into basis[col+1] fused with its CGS dots against basis[0..col]
// (per-slice partials, as launch_mdot); irfft_mdot_rows: the partial rows it
// writes (its grid), 0 where the fused kernel does not apply
int irfft_mdot_rows(const vv_ctx *c, bool fp32);
// rows per thread of the LU tridiagonal solve, and the pivots it stores per
// mode (chunk-interleaved, >= nv)
int lu_rows(int nv);
int lu_pivots_per_mode(int nv);
hipError_t launch_irfft_mdot(vv_ctx *c, const double2 *spec, int col, const int32_t *done);
hipError_t launch_irfft_mdot32(vv_ctx *c, const float2 *spec, int col, const int32_t *done);
// potential preconditioner: field (nv rows of nx) <-> transposed (nx rows of nv)
// reciprocal LU pivots of the kinetic T_k into rp64 (FP64) and / or rp32 (rounded)
hipError_t launch_tridiag_factor(vv_ctx *c, double2 *rp64, float2 *rp32);
hipError_t launch_transpose(vv_ctx *c, const double *in, double *out, bool back, const int32_t *done);
hipError_t launch_transpose32(vv_ctx *c, const float *in, float *out, bool back, const int32_t *done);
hipError_t launch_potential_bands(vv_ctx *c, const double *phi);
// krylov.hip
hipError_t launch_gmres_init(vv_ctx *c, double rtol, double atol, double dtol, int max_it);
hipError_t launch_norm_start(vv_ctx *c);                  // cycle start on basis[0]
int cgs_blocks_max(int64_t N);                            // partial rows the CGS kernels may use
int cgs_ept(const vv_ctx *c, int64_t nvv);              // 16-byte vectors per thread of a CGS launch
hipError_t launch_mdot(vv_ctx *c, int col);               // CGS dots of basis[col+1] vs 0..col
// CGS update of basis[it+1] + Givens; dot_rows: partial rows of the dots
// launch before it (0: launch_mdot's grid)
hipError_t launch_maxpy_norm(vv_ctx *c, int it, int dot_rows = 0);
hipError_t launch_build_solution(vv_ctx *c, double *x);  // x += V y
hipError_t launch_start32(vv_ctx *c, const double *r);    // FP32 cycle start: V_0 = (float) r
hipError_t launch_mdot32(vv_ctx *c, int col);
hipError_t launch_maxpy_norm32(vv_ctx *c, int it, int dot_rows = 0);
hipError_t launch_build_solution32(vv_ctx *c, double *x);  // x (FP64) += V y (FP32 basis)
hipError_t launch_residual_restart(vv_ctx *c, const double *b, double *y);  // y = b - y (B fields)
hipError_t launch_norms(vv_ctx *c, const double *x, double *out_dev);       // ||x_b|| of the B fields
hipError_t launch_norm2(vv_ctx *c, const double *x, int64_t n, double *out_dev);
hipError_t launch_dot(vv_ctx *c, const double *x, const double *y, int64_t n, double *out_dev);
hipError_t launch_axpy(vv_ctx *c, double alpha, const double *x, double *y, int64_t n);
hipError_t launch_axpy_cond(vv_ctx *c, double alpha, const double *x, double *y);  // per instance, if reason > 0
hipError_t launch_copy(vv_ctx *c, const double *x, double *y, int64_t n);
hipError_t launch_zero(vv_ctx *c, double *y, int64_t n);
hipError_t launch_xadd(vv_ctx *c, double *y, const double *w, double alpha);  // y[b][j][i] += alpha w[b][i]
// moments.hip
// density partials (DCH x nx per instance) into npart, or the context's own
// buffer when npart is null
hipError_t launch_density_partial(vv_ctx *c, const double *f, double *npart);
hipError_t launch_density(vv_ctx *c, const double *f, double *n);
hipError_t launch_velocity_moments(vv_ctx *c, const double *f, double *u, double *e);
// Poisson pipeline; npart / work (3 nx + 1 per instance) default to the
// context's buffers when null; the residual norm of instance b goes to
// pres[b * pres_stride]
hipError_t launch_poisson(vv_ctx *c, const double *n_in, double *n_out, double *phi, double *h1, double *pres,
                          int from_part, const double *npart = nullptr, double *work = nullptr,
                          int pres_stride = 16);
hipError_t launch_poisson_mult(vv_ctx *c, const double *x, const double *n, double *y);
hipError_t launch_pot_to_ham(vv_ctx *c, const double *phi, double *h1);
hipError_t launch_invariants(vv_ctx *c, const double *f, const double *h0, const double *h1,
                             const double *h2, const double *phi, double *out_dev);

// Launch a timed kernel: when the profiler armed c->prof_a / c->prof_b, the
// events are attached to this kernel's own dispatch packet (start / end
// timestamps of the kernel, as a kernel trace reports them).
template <typename K, typename... Args>
inline void launch_timed(vv_ctx *c, K kernel, dim3 grid, dim3 block, uint32_t shmem, Args... args)
{
    if (c->prof_a != nullptr) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, c->stream, c->prof_a, c->prof_b, 0u, args...);
        c->prof_a = c->prof_b = nullptr;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, c->stream, args...);
    }
}

// element stride between the spectra of consecutive instances
inline int64_t spec_stride(const vv_ctx *c) { return (int64_t)c->nmodes * c->fft_rows; }

}  // namespace vv

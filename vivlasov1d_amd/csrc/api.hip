// C ABI of libvv1d.so: context set-up, the operator entry points and the
// native GMRES / SNES(ksponly) drivers, for one instance or a batch of
// instances that share nx, nv, the v-grid and dt.  See include/vv1d.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vv1d_internal.h"

using namespace vv;

namespace {

thread_local std::string g_create_error;

enum ProfKind {
    P_JAC = 0, P_RES, P_RFFT, P_TRI, P_IRFFT, P_MDOT, P_MAXPY,
    P_JAC32, P_RFFT32, P_TRI32, P_IRFFT32, P_MDOT32, P_MAXPY32, P_IRFFT_MDOT, P_IRFFT_MDOT32, P_JAC_RFFT, P_JAC_RFFT32,
    P_NKINDS
};
const char *kProfNames[P_NKINDS] = {"jacobian",   "residual", "rfft",      "tridiag", "irfft",
                                    "mdot",       "maxpy_norm", "jacobian32", "rfft32",  "tridiag32",
                                    "irfft32",    "mdot32",   "maxpy_norm32", "irfft_mdot", "irfft_mdot32",
                                    "jac_rfft",   "jac_rfft32"};

// device scalar regions (doubles): per-instance norms | Poisson residuals
// (16 apart) | invariants (6 per instance)
inline double *ds_norms(vv_ctx *c) { return c->dscalar; }
inline double *ds_pres(vv_ctx *c) { return c->dscalar + c->B; }
inline double *ds_inv(vv_ctx *c) { return c->dscalar + 17 * (size_t)c->B; }
inline size_t ds_count(int B) { return 24 * (size_t)B + 16; }

struct ProfRec {
    int kind;
    int tag;          // Arnoldi iteration within the cycle, -1 outside GMRES
    double bytes;     // algorithmic HBM bytes of this launch
    hipEvent_t a, b;
};

struct Prof {
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    // time one launch in every `every` of each enabled kind: the events put
    // idle time of their own around a timed dispatch (several microseconds),
    // so timing every launch would slow the run it measures
    int every = 1;
    int64_t seen[P_NKINDS] = {};
};

// one profiler per context, kept outside the C struct
Prof *prof_of(vv_ctx *c) { return reinterpret_cast<Prof *>(c->prof_state); }

hipEvent_t prof_event(Prof *p)
{
    if (!p->pool.empty()) {
        hipEvent_t e = p->pool.back();
        p->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct ProfScope {
    vv_ctx *c;
    int kind;
    int tag;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    // The events are attached to the main kernel's own dispatch
    // (hipExtLaunchKernelGGL in launch_timed), so they time the kernel itself.
    // Nothing is timed while a graph is being captured.
    ProfScope(vv_ctx *cc, int k, double by, int tg = -1) : c(cc), kind(k), tag(tg), bytes(by)
    {
        if (((c->prof >> kind) & 1) && !c->capturing && prof_of(c)->seen[kind]++ % prof_of(c)->every == 0) {
            Prof *p = prof_of(c);
            a = prof_event(p);
            b = prof_event(p);
            if (a && b) {
                c->prof_a = a;
                c->prof_b = b;
            }
        }
    }
    ~ProfScope()
    {
        c->prof_a = c->prof_b = nullptr;
        if (a && b) prof_of(c)->recs.push_back({kind, tag, bytes, a, b});
    }
};

int fail(vv_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int hipfail(vv_ctx *c, hipError_t e, const char *where)
{
    return fail(c, VV_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define VV_CHECK(call, where)                                  \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return hipfail(c, _e, where);    \
    } while (0)

#define VV_PLAUNCH(kind, bytes, tag, call, where)              \
    do {                                                       \
        ProfScope _ps(c, kind, (double)(bytes), tag);          \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return hipfail(c, _e, where);    \
    } while (0)

// Device allocation owned by the context; its size is added to the context's
// ledger, which vv_scratch_bytes reports (buffers allocated on first use
// included).
template <typename T>
hipError_t dalloc(vv_ctx *c, T **p, size_t count)
{
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    if (e == hipSuccess) c->dev_bytes += (int64_t)bytes;
    return e;
}

// count doubles from device memory into the mapped pinned staging buffer
__global__ void scalars_to_host_kernel(const double *__restrict__ src, double *dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// the statuses of the B instances into a mapped pinned slot, one lane per
// instance (per-field stores)
__global__ void status_to_host_kernel(const GmresState *__restrict__ g, GmresStatus *dst, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const GmresStatus st = g[b].st;
    dst[b].it = st.it;
    dst[b].its = st.its;
    dst[b].reason = st.reason;
    dst[b].done = st.done;
    dst[b].active = st.active;
    dst[b].res = st.res;
    dst[b].rnorm0 = st.rnorm0;
}

// the instance mask from its mapped pinned staging into device memory
__global__ void mask_from_host_kernel(const int32_t *src, int32_t *__restrict__ dst, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) dst[b] = src[b];
}

// the five 1D factors of H in one launch (n = nx B for the x-factors, nv for h0)
struct FactorCopy {
    const double *src[5];
    double *dst[5];
    int n[5];
};
__global__ void factors_copy_kernel(FactorCopy a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 5; ++q)
        if (i < a.n[q]) a.dst[q][i] = a.src[q][i];
}

// count doubles from device to host: a small kernel writes them into the
// mapped pinned buffer (no copy-engine dispatch and its host-side set-up),
// then the stream is drained
int read_scalars(vv_ctx *c, const double *dev, size_t count, double *out)
{
    const int n = (int)count;
    hipLaunchKernelGGL(scalars_to_host_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, dev,
                       c->host_scalars_dev, n);
    VV_CHECK(hipGetLastError(), "copy scalars");
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    for (size_t q = 0; q < count; ++q) out[q] = c->host_scalars[q];
    return VV_OK;
}

int check_ctx(vv_ctx *c)
{
    if (!c) return fail(nullptr, VV_ERR_INVALID, "null context");
    return VV_OK;
}

// algorithmic HBM bytes per launch (compulsory reads + writes of each kernel,
// all instances of the context)
double b_field(vv_ctx *c) { return 8.0 * (double)c->N * c->B; }
double b_spec(vv_ctx *c) { return 16.0 * (double)c->nmodes * (double)c->fft_rows * c->B; }

// Arnoldi iterations enqueued past the end of the cycle return at once: their
// launches keep their (short) time but move no algorithmic bytes.
void prof_mark_idle(vv_ctx *c, size_t from, int done_its)
{
    if (!c->prof) return;
    Prof *p = prof_of(c);
    for (size_t q = from; q < p->recs.size(); ++q)
        if (p->recs[q].tag >= done_its) p->recs[q].bytes = 0.0;
}

// y = P^-1 x (spectrum in c->spec); done: optional early-exit flag
// y = P^-1 x (tr: P^-T x, the mode blocks T_k^H) of the tensor-product
// preconditioner: rFFT, one tridiagonal system per mode, irFFT.  The
// potential form transforms along v, so its field is transposed first and
// back after (x and y may alias: the transforms go through the spectrum and
// the transposed scratch).  it: the Arnoldi step (profile record) or -1.
int pc_apply(vv_ctx *c, const double *x, double *y, const int32_t *done, int it, bool tr)
{
    const double F = b_field(c), S = b_spec(c);
    const bool pot = c->tb != nullptr;
    const double *src = x;
    double *dst = y;
    if (pot) {
        VV_CHECK(launch_transpose(c, x, c->tb, false, done), "transpose");
        src = dst = c->tb;
    }
    VV_PLAUNCH(P_RFFT, F + S, it, launch_rfft(c, src, c->spec, done), "rfft");
    if (tr) {
        VV_CHECK(launch_tridiag_t(c, c->spec, done), "tridiag (transposed)");
    } else {
        VV_PLAUNCH(P_TRI, 2 * S, it, launch_tridiag_k(c, c->spec, done), "tridiag");
    }
    VV_PLAUNCH(P_IRFFT, S + F, it, launch_irfft(c, c->spec, dst, done), "irfft");
    if (pot) VV_CHECK(launch_transpose(c, c->tb, y, true, done), "transpose");
    return VV_OK;
}

int precond(vv_ctx *c, const double *x, double *y, const int32_t *done)
{
    if (c->p.precond == VV_PRECOND_NONE) {
        if (x != y) VV_CHECK(launch_copy(c, x, y, c->N * c->B), "copy");
        return VV_OK;
    }
    return pc_apply(c, x, y, done, -1, false);
}

// y = P^-T x: the transposed preconditioner (mode blocks T_k^H)
int precond_t(vv_ctx *c, const double *x, double *y, const int32_t *done)
{
    return pc_apply(c, x, y, done, -1, true);
}

int ensure_linear(vv_ctx *c)
{
    if (c->T2) return VV_OK;
    VV_CHECK(dalloc(c, &c->T2, (size_t)c->N * c->B), "alloc linearisation scratch");
    VV_CHECK(dalloc(c, &c->xs, (size_t)4 * c->p.nx * c->B), "alloc x scratch");
    return VV_OK;
}

// x scratch regions of B nx each: n | phi | h1
inline double *xs_h1(vv_ctx *c) { return c->xs + 2 * (size_t)c->p.nx * c->B; }

// h1[g] of a field g: the h1 output of the moments pipeline,
// (I - mean) C (-q) (I - mean) hv sum_j g (C: the symmetric mean-free inverse
// Laplacian), into xs_h1.  It is the derivative of h1 with respect to f, and,
// composed with the broadcast over v-rows, its own transpose.
int h1_of(vv_ctx *c, const double *g)
{
    const size_t bx = (size_t)c->p.nx * c->B;
    VV_CHECK(launch_density_partial(c, g, nullptr), "density");
    VV_CHECK(launch_poisson(c, nullptr, c->xs, c->xs + bx, c->xs + 2 * bx, nullptr, 1), "poisson");
    return VV_OK;
}

// The linearisation of the implicit-midpoint residual
//   G(f1, f0) = (f1 - f0)/dt + R {(f0 + f1)/2, H(f0, f1)},
//   H = h0 + h1[f1]/2 + h1[f0]/2 (+ h2 terms, constant),
// about (f0, f1) with f_bar = (f0 + f1)/2:
//   A  = dG/df1 =  I/dt + 1/2 R {., H} + 1/2 R {f_bar, h1[.]}
//   A0 = dG/df0 = -I/dt + 1/2 R {., H} + 1/2 R {f_bar, h1[.]}
// (R keeps rows 2..nv-3; the coupling term is dropped when fbar is null).
// Transposes, from the skew-symmetry of the stencils in their first
// argument and their antisymmetry in the two arguments:
//   A^T u  =  u/dt - 1/2 {u~, H} + 1/2 h1[{u~, f_bar}]   (all rows; u~ = R^T u)
//   A0^T u = -u/dt - 1/2 {u~, H} + 1/2 h1[{u~, f_bar}]
// with h1[g] (x-only) broadcast over the rows.  linear_apply forms
//   forward:    y = alpha u + beta R {u, H} + gamma R {f_bar, h1[u]}
//   transposed: y = alpha u + beta {u~, H} + gamma h1[{u~, f_bar}]
int linear_apply(vv_ctx *c, const double *u, const double *uscale, const double *fbar, double *y, double alpha,
                 double beta, double gamma, bool transposed)
{
    if (int r = ensure_linear(c)) return r;
    const double *h1 = xs_h1(c);
    if (transposed) {
        VV_CHECK(launch_bracket_t(c, u, uscale, nullptr, nullptr, y, alpha, beta), "bracket (transposed)");
        if (fbar && gamma != 0.0) {
            VV_CHECK(launch_bracket_t(c, u, uscale, fbar, nullptr, c->T2, 0.0, 1.0), "coupling bracket");
            if (int r = h1_of(c, c->T2)) return r;
            VV_CHECK(launch_xadd(c, y, h1, gamma), "coupling");
        }
    } else {
        if (uscale) return fail(c, VV_ERR_INVALID, "internal: scaled input of the forward linearisation");
        VV_CHECK(launch_zero(c, y, c->N * c->B), "zero");
        VV_CHECK(launch_axpy(c, alpha, u, y, c->N * c->B), "axpy");
        VV_CHECK(launch_bracket_add(c, u, y, beta), "bracket");
        if (fbar && gamma != 0.0) {
            if (int r = h1_of(c, u)) return r;
            VV_CHECK(launch_bracket_gen(c, fbar, nullptr, h1, y, gamma), "coupling bracket");
        }
    }
    return VV_OK;
}

// ---- dissipation terms (PETScNLVlasovMP.pyx:38-39, 56-58) ------------------
bool diss_on(const vv_ctx *c) { return c->diss_kind != VV_DISSIPATION_NONE; }
bool diss_coll(const vv_ctx *c)
{
    return c->diss_kind == VV_DISSIPATION_COLLT || c->diss_kind == VV_DISSIPATION_COLLE;
}

int need_moments(vv_ctx *c)
{
    const double *const *m = c->mom;
    if (!m[0] || !m[1] || !m[3] || !m[4] || !m[5] || !m[7])
        return fail(c, VV_ERR_INVALID, "collision operator: set_moments first");
    return VV_OK;
}

// The terms function_solver adds after B(Fave, H): the collision operator at
// f with the current moments and at fh with the history moments (each with
// factor 1/2), or the double bracket of Fave = .5 fh + .5 f with factor 1.
// Fave and b = {Fave, H_ave} stay in the context for the Jacobian's double
// bracket, as the reference's self.Fave does.
int diss_function(vv_ctx *c, const double *f, const double *fh, double *y)
{
    if (!diss_on(c)) return VV_OK;
    if (diss_coll(c)) {
        if (int r = need_moments(c)) return r;
        const double *const *m = c->mom;
        VV_CHECK(launch_collision(c, f, nullptr, y, m[0], m[1], m[3], 0.5, nullptr), "collision");
        VV_CHECK(launch_collision(c, fh, nullptr, y, m[4], m[5], m[7], 0.5, nullptr), "collision");
        return VV_OK;
    }
    VV_CHECK(launch_average(c, f, fh, c->fave), "average");
    VV_CHECK(launch_bracket_acc(c, c->fave, nullptr, nullptr, c->db1, 1.0, true, nullptr), "double bracket");
    c->db1_valid = true;
    const double fac = -c->coll_freq * 1.0;
    if (c->diss_kind == VV_DISSIPATION_FFH)        // {Fave, b}  (DoubleBracket.pyx:80-82)
        VV_CHECK(launch_bracket_acc(c, c->fave, nullptr, c->db1, y, fac, false, nullptr), "double bracket");
    else                                            // {b, H_ave}  (DoubleBracket.pyx:95-97)
        VV_CHECK(launch_bracket_acc(c, c->db1, nullptr, nullptr, y, fac, false, nullptr), "double bracket");
    return VV_OK;
}

// The terms jacobian_solver adds after .5 B(u, H) for u = *us * x (us may be
// null): the collision operator with the current moments and factor 1/2, or
// the double bracket's Jacobian with factor 1/2 (DoubleBracket.pyx:69-76,
// 88-91).  done: early exit of kernels enqueued past a cycle's convergence.
int diss_jacobian(vv_ctx *c, const double *x, const double *us, double *y, const int32_t *done)
{
    if (!diss_on(c)) return VV_OK;
    if (diss_coll(c)) {
        if (int r = need_moments(c)) return r;
        VV_CHECK(launch_collision(c, x, us, y, c->mom[0], c->mom[1], c->mom[3], 0.5, done), "collision");
        return VV_OK;
    }
    const double fac = -c->coll_freq * 0.5;
    if (c->diss_kind == VV_DISSIPATION_FFH) {
        if (!c->db1_valid) {
            VV_CHECK(launch_bracket_acc(c, c->fave, nullptr, nullptr, c->db1, 1.0, true, nullptr), "double bracket");
            c->db1_valid = true;
        }
        VV_CHECK(launch_bracket_acc(c, x, us, c->db1, y, fac, false, done), "double bracket");      // {u, {Fave, H}}
        VV_CHECK(launch_bracket_acc(c, x, us, nullptr, c->db2, 1.0, true, done), "double bracket");  // {u, H}
        VV_CHECK(launch_bracket_acc(c, c->fave, nullptr, c->db2, y, fac, false, done), "double bracket");
    } else {
        VV_CHECK(launch_bracket_acc(c, x, us, nullptr, c->db2, 1.0, true, done), "double bracket");  // {u, H}
        VV_CHECK(launch_bracket_acc(c, c->db2, nullptr, nullptr, y, fac, false, done), "double bracket");
    }
    return VV_OK;
}

int no_diss(vv_ctx *c, const char *what)
{
    if (!diss_on(c)) return VV_OK;
    return fail(c, VV_ERR_UNSUPPORTED, std::string(what) + " does not take a dissipation term");
}

// the preconditioner of the GMRES cycle: P^-1, or P^-T for the adjoint
int apply_pc(vv_ctx *c, const double *x, double *y, const int32_t *done)
{
    if (c->op_mode == 1) return precond_t(c, x, y, done);
    return precond(c, x, y, done);
}

// one Arnoldi step of GMRES on P^-1 J: V_{it+1} from V_it (every instance)
int arnoldi_step(vv_ctx *c, int it)
{
    const int32_t *done = &c->gs->st.done;
    const size_t vs = (size_t)c->vs;
    double *v = c->basis + (size_t)it * vs;
    double *w = c->basis + (size_t)(it + 1) * vs;
    const double F = b_field(c), S = b_spec(c);
    if (c->op_mode != 0) {
        if (int r = linear_apply(c, v, &c->gs->scale[it], c->op_fbar, c->X, c->ht_inv, -0.5, 0.5, true)) return r;
        if (int r = apply_pc(c, c->X, w, done)) return r;
    } else if (c->p.precond == VV_PRECOND_NONE) {
        // no preconditioner (the reference's preconditioner_type None): J v
        // goes straight into the next basis vector
        VV_PLAUNCH(P_JAC, 2 * F, it, launch_jacobian(c, v, &c->gs->scale[it], w), "jacobian");
        if (int r = diss_jacobian(c, v, &c->gs->scale[it], w, done)) return r;
    } else if (const int rows = irfft_mdot_rows(c, false)) {
        // P^-1 J v with the irFFT and the CGS dots of w in one launch; J v
        // and the rFFT in one launch too where no dissipation term is added
        if (jac_rfft_rows(c) && !diss_on(c)) {
            VV_PLAUNCH(P_JAC_RFFT, F + S, it, launch_jac_rfft(c, v, &c->gs->scale[it], c->spec, done), "jac_rfft");
        } else {
            VV_PLAUNCH(P_JAC, 2 * F, it, launch_jacobian(c, v, &c->gs->scale[it], c->X), "jacobian");
            if (int r = diss_jacobian(c, v, &c->gs->scale[it], c->X, done)) return r;
            VV_PLAUNCH(P_RFFT, F + S, it, launch_rfft(c, c->X, c->spec, done), "rfft");
        }
        VV_PLAUNCH(P_TRI, 2 * S, it, launch_tridiag_k(c, c->spec, done), "tridiag");
        VV_PLAUNCH(P_IRFFT_MDOT, S + F * (it + 2), it, launch_irfft_mdot(c, c->spec, it, done), "irfft_mdot");
        VV_PLAUNCH(P_MAXPY, F * (it + 3), it, launch_maxpy_norm(c, it, rows), "maxpy_norm");
        return VV_OK;
    } else if (jac_rfft_rows(c) && !diss_on(c)) {
        // the kinetic P^-1 J v with J v fused into the rFFT; irFFT and dots apart
        VV_PLAUNCH(P_JAC_RFFT, F + S, it, launch_jac_rfft(c, v, &c->gs->scale[it], c->spec, done), "jac_rfft");
        VV_PLAUNCH(P_TRI, 2 * S, it, launch_tridiag_k(c, c->spec, done), "tridiag");
        VV_PLAUNCH(P_IRFFT, S + F, it, launch_irfft(c, c->spec, w, done), "irfft");
    } else {
        VV_PLAUNCH(P_JAC, 2 * F, it, launch_jacobian(c, v, &c->gs->scale[it], c->X), "jacobian");
        if (int r = diss_jacobian(c, v, &c->gs->scale[it], c->X, done)) return r;
        if (int r = pc_apply(c, c->X, w, done, it, false)) return r;
    }
    VV_PLAUNCH(P_MDOT, F * (it + 2), it, launch_mdot(c, it), "mdot");
    VV_PLAUNCH(P_MAXPY, F * (it + 3), it, launch_maxpy_norm(c, it), "maxpy_norm");
    return VV_OK;
}

// y = P^-1 x in FP32 (complex64 spectrum in c->spec)
// FP32 P^-1 (the mixed solver's cycle), as pc_apply
// FP32 pivots of the kinetic solve, formed on first FP32 use
int pc_prepare32(vv_ctx *c)
{
    if (c->lu64 && !c->lu32) {
        VV_CHECK(dalloc(c, &c->lu32, (size_t)c->nmodes * lu_pivots_per_mode(c->fft_rows) * c->B),
                 "alloc FP32 pivots");
        VV_CHECK(launch_tridiag_factor(c, nullptr, c->lu32), "FP32 pivots");
    }
    return VV_OK;
}

int pc_apply32(vv_ctx *c, const float *x, float *y, const int32_t *done, int it)
{
    if (int r = pc_prepare32(c)) return r;
    float2 *spec = reinterpret_cast<float2 *>(c->spec);
    const double F = 4.0 * (double)c->N * c->B, S = 0.5 * b_spec(c);
    const bool pot = c->tb != nullptr;
    float *tb = reinterpret_cast<float *>(c->tb);
    const float *src = x;
    float *dst = y;
    if (pot) {
        VV_CHECK(launch_transpose32(c, x, tb, false, done), "transpose32");
        src = dst = tb;
    }
    VV_PLAUNCH(P_RFFT32, F + S, it, launch_rfft32(c, src, spec, done), "rfft32");
    VV_PLAUNCH(P_TRI32, 2 * S, it, launch_tridiag32(c, spec, done), "tridiag32");
    VV_PLAUNCH(P_IRFFT32, S + F, it, launch_irfft32(c, spec, dst, done), "irfft32");
    if (pot) VV_CHECK(launch_transpose32(c, tb, y, true, done), "transpose32");
    return VV_OK;
}

int precond32(vv_ctx *c, const float *x, float *y, const int32_t *done)
{
    return pc_apply32(c, x, y, done, -1);
}

// one Arnoldi step of the FP32 cycle: V_{it+1} = P32^-1 J32 v_it, CGS in FP64
// arithmetic on the FP32 basis
int arnoldi_step32(vv_ctx *c, int it)
{
    const int32_t *done = &c->gs->st.done;
    const size_t vs = (size_t)c->vs;
    float *v = c->basis32 + (size_t)it * vs;
    float *w = c->basis32 + (size_t)(it + 1) * vs;
    float *X = reinterpret_cast<float *>(c->X);
    const double F = 4.0 * (double)c->N * c->B;
    float *ju = c->p.precond == VV_PRECOND_NONE ? w : X;
    const int rows = c->p.precond != VV_PRECOND_NONE ? irfft_mdot_rows(c, true) : 0;
    // J u and its rFFT in one launch where the fused kernels apply and no
    // dissipation term is added
    const bool fuse_ju = rows && jac_rfft_rows(c) && !diss_on(c);
    if (!fuse_ju) {
        VV_PLAUNCH(P_JAC32, 2 * F, it, launch_jacobian32(c, v, &c->gs->scale[it], ju), "jacobian32");
    }
    if (diss_on(c)) {
        // the dissipation term in FP64 on the widened vector, folded into J u
        VV_CHECK(launch_widen(c, v, c->w64), "widen");
        VV_CHECK(launch_zero(c, c->d64, c->N * c->B), "zero");
        if (int r = diss_jacobian(c, c->w64, &c->gs->scale[it], c->d64, done)) return r;
        VV_CHECK(launch_narrow_add(c, ju, c->d64), "narrow add");
    }
    if (rows) {
        // P32^-1 with the irFFT and the CGS dots of w in one launch
        if (int r = pc_prepare32(c)) return r;
        float2 *spec = reinterpret_cast<float2 *>(c->spec);
        const double S = 0.5 * b_spec(c);
        if (fuse_ju) {
            VV_PLAUNCH(P_JAC_RFFT32, F + S, it, launch_jac_rfft32(c, v, &c->gs->scale[it], spec, done), "jac_rfft32");
        } else {
            VV_PLAUNCH(P_RFFT32, F + S, it, launch_rfft32(c, X, spec, done), "rfft32");
        }
        VV_PLAUNCH(P_TRI32, 2 * S, it, launch_tridiag32(c, spec, done), "tridiag32");
        VV_PLAUNCH(P_IRFFT_MDOT32, S + F * (it + 2), it, launch_irfft_mdot32(c, spec, it, done), "irfft_mdot32");
        VV_PLAUNCH(P_MAXPY32, F * (it + 3), it, launch_maxpy_norm32(c, it, rows), "maxpy_norm32");
        return VV_OK;
    }
    if (c->p.precond != VV_PRECOND_NONE)
        if (int r = pc_apply32(c, X, w, done, it)) return r;
    VV_PLAUNCH(P_MDOT32, F * (it + 2), it, launch_mdot32(c, it), "mdot32");
    VV_PLAUNCH(P_MAXPY32, F * (it + 3), it, launch_maxpy_norm32(c, it), "maxpy_norm32");
    return VV_OK;
}

GmresStatus *status_slot(vv_ctx *c, int slot) { return c->host_status + (size_t)slot * c->B; }

bool all_done(const GmresStatus *s, int B)
{
    for (int b = 0; b < B; ++b)
        if (!s[b].done) return false;
    return true;
}

int max_it_of(const GmresStatus *s, int B)
{
    int m = 0;
    for (int b = 0; b < B; ++b) m = s[b].it > m ? s[b].it : m;
    return m;
}

// the device statuses into host slot `slot` by a kernel writing the mapped
// pinned memory (no copy-engine dispatch)
int copy_status(vv_ctx *c, int slot)
{
    hipLaunchKernelGGL(status_to_host_kernel, dim3((c->B + 255) / 256), dim3(256), 0, c->stream, c->gs,
                       c->host_status_dev + (size_t)slot * c->B, c->B);
    VV_CHECK(hipGetLastError(), "status copy");
    return VV_OK;
}

// Arnoldi steps of one GMRES cycle, polling the device status every two
// steps (the status copy of step k is read while steps k+1, k+2 run); returns
// once the cycle has finished on the device for every instance, with the
// statuses in host slot 0.  Steps enqueued past the end of an instance's
// cycle return at once on the device.  build: x += V y is enqueued before
// the final drain (FP64 x, from the FP32 basis when fp32); update: then
// f -= x for instances whose solve converged (SNESSolve_KSPONLY's step), so
// one drain ends the solve.
int run_cycle(vv_ctx *c, bool fp32, double *build = nullptr, double *update = nullptr)
{
    const int max_k = c->p.restart;
    const int chunk = 2;
    const size_t prof_from = c->prof ? prof_of(c)->recs.size() : 0;
    int it = 0, pend = -1, slot = 0;
    bool stop = false;
    while (it < max_k && !stop) {
        const int end = it + chunk < max_k ? it + chunk : max_k;
        for (; it < end; ++it) {
            // the chunk's last Hessenberg update stores the statuses into the
            // host slot itself (no copy dispatch between Arnoldi steps)
            if (it == end - 1) c->status_out = c->host_status_dev + (size_t)slot * c->B;
            int r = fp32 ? arnoldi_step32(c, it) : arnoldi_step(c, it);
            c->status_out = nullptr;
            if (r) return r;
        }
        VV_CHECK(hipEventRecord(c->ev[slot], c->stream), "event");
        if (pend >= 0) {
            VV_CHECK(hipEventSynchronize(c->ev[pend]), "event sync");
            if (all_done(status_slot(c, pend), c->B)) stop = true;
        }
        pend = slot;
        slot ^= 1;
    }
    if (build) {
        if (fp32) VV_CHECK(launch_build_solution32(c, build), "build solution");
        else VV_CHECK(launch_build_solution(c, build), "build solution");
        if (update) VV_CHECK(launch_axpy_cond(c, -1.0, build, update), "update");
    }
    if (int r = copy_status(c, 0)) return r;
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    prof_mark_idle(c, prof_from, max_it_of(status_slot(c, 0), c->B));
    return VV_OK;
}

// V_0 = b (per instance) and the cycle start
int cycle_start_from(vv_ctx *c, const double *b)
{
    VV_CHECK(launch_copy(c, b, c->basis, c->N * c->B), "copy b");
    VV_CHECK(launch_norm_start(c), "cycle start");
    return VV_OK;
}

void fill_results(vv_ctx *c, vv_ksp_result *res)
{
    if (!res) return;
    const GmresStatus *st = status_slot(c, 0);
    for (int b = 0; b < c->B; ++b) {
        res[b].its = st[b].its;
        res[b].reason = st[b].reason;
        res[b].rnorm = st[b].res;
        res[b].rnorm0 = st[b].rnorm0;
    }
}

// Cycles after the first: while some instance still iterates, restart from
// r = b - (P^-1 J) x (KSPInitialResidual) and run another cycle.  Instances
// that finished stay done (their cycle start keeps done set).
int restart_cycles(vv_ctx *c, const double *b, double *x)
{
    for (;;) {
        const GmresStatus *st = status_slot(c, 0);
        bool more = false;
        for (int q = 0; q < c->B; ++q)
            if (st[q].active && st[q].reason == 0) more = true;
        if (!more) return VV_OK;
        if (c->op_mode != 0) {
            if (int r = linear_apply(c, x, nullptr, c->op_fbar, c->X, c->ht_inv, -0.5, 0.5, true)) return r;
        } else {
            VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, x, nullptr, c->X), "jacobian");
            if (int r = diss_jacobian(c, x, nullptr, c->X, nullptr)) return r;
        }
        if (int r = apply_pc(c, c->X, c->basis, nullptr)) return r;
        VV_CHECK(launch_residual_restart(c, b, c->basis), "restart residual");
        VV_CHECK(launch_norm_start(c), "cycle start");
        if (int r = run_cycle(c, false, x)) return r;
    }
}

int ksp_solve_impl(vv_ctx *c, const double *b, double *x, const vv_ksp_params *kp, vv_ksp_result *res)
{
    VV_CHECK(launch_zero(c, x, c->N * c->B), "zero x");
    VV_CHECK(launch_gmres_init(c, kp->rtol, kp->atol, kp->dtol, kp->max_it), "gmres init");
    if (int r = cycle_start_from(c, b)) return r;
    if (int r = run_cycle(c, false, x)) return r;
    if (int r = restart_cycles(c, b, x)) return r;
    fill_results(c, res);
    return VV_OK;
}

// The first GMRES cycle of SNESSolve_KSPONLY as one captured graph:
// b = P^-1 F(f); x = 0; GMRES init; V_0 = b; cycle start; `restart` Arnoldi
// steps (steps past an instance's convergence return at once); x += V y.
// The graph is kept per (f, fh) pair and replayed while they stay the same;
// the KSP tolerances are kernel arguments, so a change re-captures too.
int snes_graph(vv_ctx *c, double *f, const double *fh, const vv_ksp_params *kp)
{
    const bool same = c->graph_exec != nullptr && c->graph_f == f && c->graph_fh == fh &&
                      c->graph_kp[0] == kp->rtol && c->graph_kp[1] == kp->atol && c->graph_kp[2] == kp->dtol &&
                      c->graph_kp[3] == (double)kp->max_it;
    if (!same) {
        if (c->graph_exec) {
            (void)hipGraphExecDestroy(c->graph_exec);
            c->graph_exec = nullptr;
        }
        if (c->stream == nullptr) return fail(c, VV_ERR_INVALID, "graph capture needs a non-null stream");
        VV_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        c->capturing = true;
        int r = VV_OK;
        hipError_t le = hipSuccess;
        do {
            if ((le = launch_residual(c, f, fh, c->X)) != hipSuccess) break;
            if ((r = precond(c, c->X, c->rhs, nullptr))) break;
            if ((le = launch_zero(c, c->D, c->N * c->B)) != hipSuccess) break;
            if ((le = launch_gmres_init(c, kp->rtol, kp->atol, kp->dtol, kp->max_it)) != hipSuccess) break;
            if ((r = cycle_start_from(c, c->rhs))) break;
            for (int it = 0; it < c->p.restart && r == VV_OK; ++it) r = arnoldi_step(c, it);
            if (r) break;
            le = launch_build_solution(c, c->D);
        } while (0);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->capturing = false;
        if (r != VV_OK || le != hipSuccess || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (r != VV_OK) return r;
            return hipfail(c, le != hipSuccess ? le : e, "graph capture");
        }
        const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->graph_exec = nullptr;
            return hipfail(c, ei, "graph instantiate");
        }
        c->graph_f = f;
        c->graph_fh = fh;
        c->graph_kp[0] = kp->rtol;
        c->graph_kp[1] = kp->atol;
        c->graph_kp[2] = kp->dtol;
        c->graph_kp[3] = (double)kp->max_it;
    }
    VV_CHECK(hipGraphLaunch(c->graph_exec, c->stream), "graph launch");
    if (int r = copy_status(c, 0)) return r;
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    return VV_OK;
}

int snes_impl(vv_ctx *c, double *f, const double *fh, const vv_ksp_params *kp, vv_ksp_result *res)
{
    if (c->use_graph) {
        if (int r = no_diss(c, "graph replay")) return r;
        if (int r = snes_graph(c, f, fh, kp)) return r;
    } else {
        VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, c->X), "residual");
        if (int r = diss_function(c, f, fh, c->X)) return r;
        if (int r = precond(c, c->X, c->rhs, nullptr)) return r;
        VV_CHECK(launch_zero(c, c->D, c->N * c->B), "zero x");
        VV_CHECK(launch_gmres_init(c, kp->rtol, kp->atol, kp->dtol, kp->max_it), "gmres init");
        if (int r = cycle_start_from(c, c->rhs)) return r;
        // one instance: the step f -= D is enqueued with the solution before
        // the drain (a no-op unless the cycle converged; a restart leaves it
        // to the update below)
        if (int r = run_cycle(c, false, c->D, c->B == 1 ? f : nullptr)) return r;
        if (c->B == 1 && status_slot(c, 0)[0].reason != 0) {
            fill_results(c, res);
            return VV_OK;
        }
    }
    if (int r = restart_cycles(c, c->rhs, c->D)) return r;
    // SNESSolve_KSPONLY takes the step unless the linear solve diverged
    VV_CHECK(launch_axpy_cond(c, -1.0, c->D, f), "update");
    fill_results(c, res);
    return VV_OK;
}

// lazily allocated buffers of the mixed-precision solver (FP32 basis, FP64 residual)
int ensure_mixed(vv_ctx *c)
{
    if (diss_on(c) && !c->w64) {
        VV_CHECK(dalloc(c, &c->w64, (size_t)c->N * c->B), "alloc widened Krylov vector");
        VV_CHECK(dalloc(c, &c->d64, (size_t)c->N * c->B), "alloc dissipation term");
    }
    if (c->basis32) return VV_OK;
    VV_CHECK(dalloc(c, &c->basis32, (size_t)(c->p.restart + 1) * c->vs), "alloc FP32 basis");
    VV_CHECK(dalloc(c, &c->R, (size_t)c->N * c->B), "alloc residual");
    return VV_OK;
}

// Mixed-precision solve of (P^-1 J) x = b: iterative refinement whose
// correction equations are solved by FP32 GMRES cycles (FP32 J u, FP32 P^-1,
// FP32 basis; Hessenberg, dots and norms in FP64) and whose residual
// r = b - P^-1 J x, solution x and stopping test are FP64.  The stopping test
// is the reference's (||r|| <= max(rtol ||b||, atol)) applied to the true
// FP64 preconditioned residual; every cycle restarts from that residual.
int ksp_solve_mixed_impl(vv_ctx *c, const double *b, double *x, const vv_ksp_params *kp, const vv_mixed_params *mp,
                         vv_ksp_result *res, int32_t *cycles)
{
    if (c->B != 1) return fail(c, VV_ERR_UNSUPPORTED, "the mixed-precision solve takes one instance");
    if (int r = ensure_mixed(c)) return r;
    const int64_t N = c->N;
    VV_CHECK(launch_zero(c, x, N), "zero x");
    double rnorm0 = 0.0, rnorm = 0.0;
    const double *rsrc = b;
    int its = 0, rounds = 0, reason = 0;
    if (!mp->fp64_check) {
        // one FP32 cycle with the KSP's own tolerances; its reason stands
        VV_CHECK(launch_gmres_init(c, kp->rtol, kp->atol, kp->dtol, kp->max_it), "gmres init");
        VV_CHECK(launch_start32(c, b), "cycle start");
        if (int r = run_cycle(c, true, x)) return r;
        const GmresStatus st = status_slot(c, 0)[0];
        if (st.reason != 0) {
            fill_results(c, res);
            if (cycles) *cycles = st.it > 0 ? 1 : 0;
            return VV_OK;
        }
        // the restart length ran out: FP64-checked refinement from here
        its = st.it;
        rounds = 1;
        rnorm0 = st.rnorm0;
        VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, x, nullptr, c->X), "jacobian");
        if (int r = diss_jacobian(c, x, nullptr, c->X, nullptr)) return r;
        if (int r = precond(c, c->X, c->R, nullptr)) return r;
        VV_CHECK(launch_residual_restart(c, b, c->R), "residual");
        VV_CHECK(launch_norm2(c, c->R, N, ds_norms(c)), "norm");
        if (int r = read_scalars(c, ds_norms(c), 1, &rnorm)) return r;
        rsrc = c->R;
    } else {
        VV_CHECK(launch_norm2(c, b, N, ds_norms(c)), "norm");
        if (int r = read_scalars(c, ds_norms(c), 1, &rnorm0)) return r;
        rnorm = rnorm0;
    }
    const double ttol = std::fmax(kp->rtol * rnorm0, kp->atol);
    for (;;) {
        if (!std::isfinite(rnorm)) {
            reason = VV_KSP_DIVERGED_NANORINF;
            break;
        }
        if (rnorm <= ttol) {
            reason = rnorm < kp->atol ? VV_KSP_CONVERGED_ATOL : VV_KSP_CONVERGED_RTOL;
            break;
        }
        if (rnorm >= kp->dtol * rnorm0) {
            reason = VV_KSP_DIVERGED_DTOL;
            break;
        }
        if (its >= kp->max_it || rounds >= mp->max_cycles) {
            reason = VV_KSP_DIVERGED_ITS;
            break;
        }
        // FP32 cycle on the correction equation (P^-1 J) e = r, to a relative
        // reduction of inner_rtol, or less when less already meets the target
        double rt = 0.5 * ttol / rnorm;
        if (rt < mp->inner_rtol) rt = mp->inner_rtol;
        VV_CHECK(launch_gmres_init(c, rt, 0.0, 1.0e300, kp->max_it - its), "gmres init");
        VV_CHECK(launch_start32(c, rsrc), "cycle start");
        if (int r = run_cycle(c, true, x)) return r;
        its += status_slot(c, 0)[0].it;
        ++rounds;
        // true FP64 residual r = b - P^-1 J x
        VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, x, nullptr, c->X), "jacobian");
        if (int r = diss_jacobian(c, x, nullptr, c->X, nullptr)) return r;
        if (int r = precond(c, c->X, c->R, nullptr)) return r;
        VV_CHECK(launch_residual_restart(c, b, c->R), "residual");
        VV_CHECK(launch_norm2(c, c->R, N, ds_norms(c)), "norm");
        if (int r = read_scalars(c, ds_norms(c), 1, &rnorm)) return r;
        rsrc = c->R;
    }
    if (res) {
        res->its = its;
        res->reason = reason;
        res->rnorm = rnorm;
        res->rnorm0 = rnorm0;
    }
    if (cycles) *cycles = rounds;
    return VV_OK;
}

int check_params(const vv_params *p)
{
    if (p->nx < 8 || p->nx > kMaxNx) return fail(nullptr, VV_ERR_INVALID, "nx must be in [8, 4096]");
    if (p->nv < 8 || p->nv > 4096) return fail(nullptr, VV_ERR_INVALID, "nv must be in [8, 4096]");
    if (p->bracket != VV_BRACKET_J1 && p->bracket != VV_BRACKET_J2 && p->bracket != VV_BRACKET_J4)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "bracket must be arakawaj1, arakawaj2 or arakawaj4");
    if (p->precond != VV_PRECOND_NONE && p->precond != VV_PRECOND_TENSOR_KINETIC &&
        p->precond != VV_PRECOND_TENSOR_POTENTIAL)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "preconditioner must be none, tensor_kinetic or tensor_potential");
    if (p->laplace != VV_LAPLACE_CFD2 && p->laplace != VV_LAPLACE_CFD4)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "laplace operator must be cfd2 or cfd4");
    if (p->restart < 1 || p->restart > kMaxRestart) return fail(nullptr, VV_ERR_INVALID, "restart must be in [1, 64]");
    if (!(p->hv > 0) || !(p->ht > 0)) return fail(nullptr, VV_ERR_INVALID, "hx, hv, ht must be > 0");
    if (!p->v) return fail(nullptr, VV_ERR_INVALID, "v grid missing");
    if (p->precond != VV_PRECOND_NONE && !p->eigen)
        return fail(nullptr, VV_ERR_INVALID, "preconditioner eigenvalues missing");
    return VV_OK;
}

// Scratch of vv_moments_noalloc: a file-scope device array, so that the entry
// point needs no allocation at all.  Every GPU holds its own copy until the
// process ends; its device address is looked up once per context, with the
// context's device current (vv_ctx_create), and only that address is used
// from the host.  Layout for X = nx * batch x-points: density partials
// (kDensityChunks * X doubles) | Poisson work (3 nx + 1 per instance, at most
// 4 X doubles).
constexpr int64_t kNoallocMaxX = 4096;          // nx * batch
constexpr int64_t kNoallocDoubles = (kDensityChunks + 4) * kNoallocMaxX;     // 2.125 MiB

}  // namespace

namespace vv {
__device__ __align__(256) double scratch[kNoallocDoubles];
}  // namespace vv

namespace {

int create_impl(const vv_params *p, int B, const double *hxs, void *stream, vv_ctx **out)
{
    vv_ctx *c = nullptr;
    if (!p || !out || !hxs) return fail(nullptr, VV_ERR_INVALID, "null argument");
    *out = nullptr;
    if (int r = check_params(p)) return r;
    if (B < 1 || B > 65535) return fail(nullptr, VV_ERR_INVALID, "batch must be in [1, 65535]");
    if (B > 1 && ((int64_t)p->nx * p->nv) % 4 != 0)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "a batched context needs nx * nv divisible by 4");
    for (int b = 0; b < B; ++b)
        if (!(hxs[b] > 0)) return fail(nullptr, VV_ERR_INVALID, "hx, hv, ht must be > 0");

    c = new (std::nothrow) vv_ctx();
    if (!c) return fail(nullptr, VV_ERR_INVALID, "out of host memory");
    c->p = *p;
    c->p.v = nullptr;
    c->p.eigen = nullptr;
    c->p.hx = hxs[0];
    c->B = B;
    c->fuse = true;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    const int nx = p->nx, nv = p->nv;
    c->N = (int64_t)nx * nv;
    c->vs = (c->N * B + 3) / 4 * 4;
    const bool pot = p->precond == VV_PRECOND_TENSOR_POTENTIAL;
    c->fft_n = pot ? nv : nx;
    c->fft_rows = pot ? nx : nv;
    c->nmodes = c->fft_n / 2 + 1;
    c->logn = 0;
    while ((1 << c->logn) < c->fft_n) ++c->logn;
    c->coef_stride = pot ? nx : 0;
    c->cf_lo = pot ? 1.0 : 0.5;     // PotentialFast.pyx:250-252 / KineticFast.pyx:258-260
    c->cf_mid = pot ? 2.0 : 1.0;
    c->cf_hi = pot ? 1.0 : 0.5;
    // Grid.create (Grid.pyx:40-48) and the bracket / band factors
    const double hv = p->hv;
    c->ht_inv = 1. / p->ht;
    c->hv_inv = 1. / hv;
    const double hv2 = hv * hv;
    c->hv2_inv = 1. / hv2;
    c->inst_host.resize(B);
    for (int b = 0; b < B; ++b) {
        InstConst &ic = c->inst_host[b];
        const double hx = hxs[b];
        ic.hx = hx;
        ic.hx_inv = 1. / hx;
        ic.hx2_inv = 1. / (hx * hx);
        ic.af1 = ic.hx_inv * c->hv_inv / 12.;          // PoissonBracket.pyx:88
        ic.af2 = ic.hx_inv * c->hv_inv / 24.;          // PoissonBracket.pyx:209
        ic.afpc = 0.5 * ic.hx_inv * c->hv_inv / 12.;   // KineticFast.pyx:251
    }
    c->prof_state = new Prof();

    auto bail = [&](hipError_t e, const char *where) {
        std::string m = std::string(where) + ": " + hipGetErrorString(e);
        vv_ctx_destroy(c);
        return fail(nullptr, VV_ERR_HIP, m);
    };
#define VV_CCHECK(call, where)                      \
    do {                                            \
        hipError_t _e = (call);                     \
        if (_e != hipSuccess) return bail(_e, where); \
    } while (0)

    const int nfb = fft_blocks(c), ncb = cgs_blocks_max(c->N);
    int nrows = (nfb > kReduceBlocks ? nfb : kReduceBlocks);
    if (ncb > nrows) nrows = ncb;
    c->part_stride = (int64_t)nrows * (kMaxRestart + 2);
    const size_t Bz = (size_t)B;
    VV_CCHECK(dalloc(c, &c->inst, Bz), "alloc instance constants");
    VV_CCHECK(dalloc(c, &c->active, Bz), "alloc active flags");
    VV_CCHECK(dalloc(c, &c->h0, nv), "alloc h0");
    VV_CCHECK(dalloc(c, &c->h1p, nx * Bz), "alloc h1p");
    VV_CCHECK(dalloc(c, &c->h1h, nx * Bz), "alloc h1h");
    VV_CCHECK(dalloc(c, &c->h2p, nx * Bz), "alloc h2p");
    VV_CCHECK(dalloc(c, &c->h2h, nx * Bz), "alloc h2h");
    VV_CCHECK(dalloc(c, &c->eigen, c->nmodes), "alloc eigen");
    const size_t ncoef = pot ? (size_t)nx * Bz : (size_t)nv;
    VV_CCHECK(dalloc(c, &c->gm, ncoef), "alloc gm");
    VV_CCHECK(dalloc(c, &c->g0, ncoef), "alloc g0");
    VV_CCHECK(dalloc(c, &c->gp, ncoef), "alloc gp");
    if (pot) VV_CCHECK(dalloc(c, &c->tb, c->N * Bz), "alloc transposed field");
    VV_CCHECK(dalloc(c, &c->vgrid, nv), "alloc v grid");
    VV_CCHECK(dalloc(c, &c->twiddle, c->fft_n), "alloc twiddle");
    VV_CCHECK(dalloc(c, &c->twiddle32, c->fft_n), "alloc twiddle32");
    VV_CCHECK(dalloc(c, &c->spec, (size_t)c->nmodes * c->fft_rows * Bz), "alloc spectrum");
    if (p->precond == VV_PRECOND_TENSOR_KINETIC)
        VV_CCHECK(dalloc(c, &c->lu64, (size_t)c->nmodes * lu_pivots_per_mode(c->fft_rows) * Bz), "alloc pivots");
    VV_CCHECK(dalloc(c, &c->pinv, nx * Bz), "alloc pinv");
    VV_CCHECK(dalloc(c, &c->pwork, (3 * (size_t)nx + 1) * Bz), "alloc Poisson scratch");
    VV_CCHECK(dalloc(c, &c->X, c->N * Bz), "alloc X");
    VV_CCHECK(dalloc(c, &c->rhs, c->N * Bz), "alloc rhs");
    VV_CCHECK(dalloc(c, &c->D, c->N * Bz), "alloc D");
    VV_CCHECK(dalloc(c, &c->basis, (size_t)(p->restart + 1) * c->vs), "alloc Krylov basis");
    VV_CCHECK(dalloc(c, &c->partial, (size_t)c->part_stride * Bz), "alloc partial");
    VV_CCHECK(dalloc(c, &c->dscalar, ds_count(B)), "alloc scalars");
    VV_CCHECK(dalloc(c, &c->npart, (size_t)64 * nx * Bz), "alloc density partials");
    VV_CCHECK(dalloc(c, &c->gs, Bz), "alloc gmres state");
    VV_CCHECK(hipHostMalloc(reinterpret_cast<void **>(&c->host_status), 2 * Bz * sizeof(GmresStatus)),
              "pinned status");
    VV_CCHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->host_status_dev), c->host_status, 0),
              "pinned status device address");
    VV_CCHECK(hipHostMalloc(reinterpret_cast<void **>(&c->host_scalars), ds_count(B) * sizeof(double)),
              "pinned scalars");
    VV_CCHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->host_scalars_dev), c->host_scalars, 0),
              "pinned scalars device address");
    VV_CCHECK(hipHostMalloc(reinterpret_cast<void **>(&c->host_active), Bz * sizeof(int32_t)), "pinned mask");
    VV_CCHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->host_active_dev), c->host_active, 0),
              "pinned mask device address");
    VV_CCHECK(hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming), "event");
    VV_CCHECK(hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming), "event");
    VV_CCHECK(hipMemset(c->gs, 0, Bz * sizeof(GmresState)), "zero gmres state");
    VV_CCHECK(hipMemset(c->partial, 0, (size_t)c->part_stride * Bz * sizeof(double)), "zero partials");
    VV_CCHECK(hipMemset(c->h1p, 0, nx * Bz * sizeof(double)), "zero h1p");
    VV_CCHECK(hipMemset(c->h1h, 0, nx * Bz * sizeof(double)), "zero h1h");
    VV_CCHECK(hipMemset(c->h2p, 0, nx * Bz * sizeof(double)), "zero h2p");
    VV_CCHECK(hipMemset(c->h2h, 0, nx * Bz * sizeof(double)), "zero h2h");
    VV_CCHECK(hipMemset(c->h0, 0, nv * sizeof(double)), "zero h0");

    // host tables
    // kinetic band factors of v (KineticFast.pyx:258-260); the potential
    // factors are phi differences, zero until vv_precond_update
    std::vector<double> gm(ncoef, 0.0), g0(ncoef, 0.0), gp(ncoef, 0.0);
    for (int j = 0; j < nv && !pot; ++j) {
        const double v = p->v[j];
        gm[j] = 2. * hv * v - hv2;     // KineticFast.pyx:258
        g0[j] = 4. * hv * v;           // :259
        gp[j] = 2. * hv * v + hv2;     // :260
    }
    std::vector<double2> eig(c->nmodes);
    for (int k = 0; k < c->nmodes; ++k) {
        eig[k].x = p->eigen ? p->eigen[2 * k] : 0.0;
        eig[k].y = p->eigen ? p->eigen[2 * k + 1] : 0.0;
    }
    const int nf = c->fft_n;
    std::vector<double2> tw(nf);
    std::vector<float2> tw32(nf);
    for (int m = 0; m < nf; ++m) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)nf;
        tw[m].x = (double)cosl(ang);
        tw[m].y = (double)sinl(ang);
        tw32[m].x = (float)cosl(ang);
        tw32[m].y = (float)sinl(ang);
    }
    // mean-free pseudo-inverse of the periodic Laplacian, per instance:
    // c_m = 1/nx sum_{k>0} cos(2 pi k m/nx) / lambda_k
    std::vector<long double> ct(nx), lam(nx);
    for (int q = 0; q < nx; ++q)
        ct[q] = cosl(2.0L * 3.141592653589793238462643383279502884L * (long double)q / (long double)nx);
    std::vector<double> pinv(nx * Bz);
    for (int b = 0; b < B; ++b) {
        const double hx = hxs[b];
        for (int k = 1; k < nx; ++k) {
            const long double th2 = 3.141592653589793238462643383279502884L * (long double)k / (long double)nx;
            const long double s1 = sinl(th2), s2 = sinl(2.0L * th2);
            const long double h2 = (long double)hx * (long double)hx;
            // CFD4 symbol (54 - 64 cos t + 10 cos 2t)/12 = (128 sin^2(t/2) - 20 sin^2 t)/12; CFD2: 4 sin^2(t/2)
            lam[k] = (p->laplace == VV_LAPLACE_CFD4) ? (128.0L * s1 * s1 - 20.0L * s2 * s2) / (12.0L * h2)
                                                     : (4.0L * s1 * s1) / h2;
        }
        for (int m = 0; m < nx; ++m) {
            long double s = 0.0L;
            for (int k = 1; k < nx; ++k) s += ct[(int)(((long long)k * m) % nx)] / lam[k];
            pinv[(size_t)b * nx + m] = (double)(s / (long double)nx);
        }
    }
    std::vector<int32_t> act(Bz, 1);
    VV_CCHECK(hipMemcpy(c->inst, c->inst_host.data(), Bz * sizeof(InstConst), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->active, act.data(), Bz * sizeof(int32_t), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->gm, gm.data(), ncoef * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->g0, g0.data(), ncoef * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->gp, gp.data(), ncoef * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->vgrid, p->v, nv * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->eigen, eig.data(), c->nmodes * sizeof(double2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->twiddle, tw.data(), nf * sizeof(double2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->twiddle32, tw32.data(), nf * sizeof(float2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->pinv, pinv.data(), nx * Bz * sizeof(double), hipMemcpyHostToDevice), "upload");
    // reciprocal LU pivots of the kinetic T_k (fixed for the run)
    if (c->lu64) VV_CCHECK(launch_tridiag_factor(c, c->lu64, nullptr), "pivots");
    // this device's copy of the allocation-free entry point's scratch array
    VV_CCHECK(hipGetSymbolAddress(reinterpret_cast<void **>(&c->static_scratch), HIP_SYMBOL(vv::scratch)),
              "static scratch address");
#undef VV_CCHECK
    *out = c;
    return VV_OK;
}

}  // namespace

extern "C" {

const char *vv_last_error(const vv_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

int vv_ctx_create(const vv_params *p, void *stream, vv_ctx **out)
{
    if (!p) return fail(nullptr, VV_ERR_INVALID, "null argument");
    return create_impl(p, 1, &p->hx, stream, out);
}

int vv_ctx_create_batched(const vv_params *p, int32_t batch, const double *hx, void *stream, vv_ctx **out)
{
    return create_impl(p, batch, hx, stream, out);
}

int32_t vv_batch_size(const vv_ctx *c) { return c ? c->B : 0; }

int vv_ctx_destroy(vv_ctx *c)
{
    if (!c) return VV_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    void *bufs[] = {c->inst,   c->active, c->h0,      c->h1p,     c->h1h,  c->h2p,   c->h2h,    c->eigen,
                    c->gm,     c->g0,     c->gp,      c->twiddle, c->twiddle32, c->spec, c->pinv, c->pwork,
                    c->X,      c->rhs,    c->D,       c->basis,   c->basis32, c->R,   c->partial, c->dscalar,
                    c->npart,  c->gs,     c->T2,      c->xs,   c->vgrid, c->tb,     c->fave,    c->db1,
                    c->db2,    c->ex,     c->w64,     c->d64,   c->lu64,   c->lu32};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->host_status) (void)hipHostFree(c->host_status);
    if (c->host_scalars) (void)hipHostFree(c->host_scalars);
    if (c->host_active) (void)hipHostFree(c->host_active);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    Prof *pr = prof_of(c);
    if (pr) {
        for (auto &r : pr->recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pr->pool) (void)hipEventDestroy(e);
        delete pr;
    }
    delete c;
    return VV_OK;
}

int vv_set_stream(vv_ctx *c, void *stream)
{
    if (int r = check_ctx(c)) return r;
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->stream = reinterpret_cast<hipStream_t>(stream);
    return VV_OK;
}

int vv_synchronize(vv_ctx *c)
{
    if (int r = check_ctx(c)) return r;
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    return VV_OK;
}

int64_t vv_scratch_bytes(const vv_ctx *c)
{
    return c ? c->dev_bytes : 0;
}

int vv_set_active(vv_ctx *c, const int32_t *active)
{
    if (int r = check_ctx(c)) return r;
    // stream-ordered: a kernel on the context's stream reads the mapped
    // pinned staging (no legacy-stream operation, so other streams may be
    // capturing graphs meanwhile); the drain first keeps an earlier upload
    // from seeing the new values
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    for (int b = 0; b < c->B; ++b) c->host_active[b] = active ? (active[b] ? 1 : 0) : 1;
    hipLaunchKernelGGL(mask_from_host_kernel, dim3((c->B + 255) / 256), dim3(256), 0, c->stream,
                       c->host_active_dev, c->active, c->B);
    VV_CHECK(hipGetLastError(), "upload mask");
    return VV_OK;
}

int vv_set_graph(vv_ctx *c, int enable)
{
    if (int r = check_ctx(c)) return r;
    if (enable)
        if (int r = no_diss(c, "graph replay")) return r;
    if (enable && c->stream == nullptr) return fail(c, VV_ERR_INVALID, "graph replay needs a non-null stream");
    if (!enable && c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->use_graph = enable != 0;
    return VV_OK;
}

int vv_set_tuning(vv_ctx *c, int32_t fft_rows, int32_t cgs_ept)
{
    if (int r = check_ctx(c)) return r;
    if (fft_rows != 0 && (fft_rows < 2 || (fft_rows & 1) || fft_rows * c->fft_n > 8192))
        return fail(c, VV_ERR_INVALID, "fft_rows must be 0 or even, >= 2, with fft_rows * (FFT length) <= 8192");
    if (cgs_ept != 0 && cgs_ept != 1 && cgs_ept != 2 && cgs_ept != 4 && cgs_ept != 8)
        return fail(c, VV_ERR_INVALID, "cgs_ept must be 0, 1, 2, 4 or 8");
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->tune_fft_rows = fft_rows;
    c->tune_cgs_ept = cgs_ept;
    return VV_OK;
}

int vv_set_fusion(vv_ctx *c, int enable)
{
    if (int r = check_ctx(c)) return r;
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->fuse = enable != 0;
    return VV_OK;
}

int vv_get_tuning(vv_ctx *c, int32_t *fft_rows, int32_t *cgs_ept)
{
    if (int r = check_ctx(c)) return r;
    if (fft_rows) *fft_rows = c->tune_fft_rows;
    if (cgs_ept) *cgs_ept = c->tune_cgs_ept;
    return VV_OK;
}

// Times the candidate launch shapes on this context's own buffers and keeps
// the fastest: v-rows per FFT workgroup (over the FP64 preconditioner) and
// vectors per thread of the CGS kernels (over mdot + maxpy at a mid-cycle
// basis size).  Overwrites the work buffers and the GMRES state.
int vv_autotune(vv_ctx *c, int32_t *fft_rows, int32_t *cgs_ept)
{
    if (int r = check_ctx(c)) return r;
    if (c->p.precond == VV_PRECOND_NONE)
        return fail(c, VV_ERR_UNSUPPORTED, "autotune needs a tensor-product preconditioner");
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    const int prof = c->prof;
    c->prof = 0;
    hipEvent_t e0, e1;
    VV_CHECK(hipEventCreate(&e0), "event");
    VV_CHECK(hipEventCreate(&e1), "event");
    auto timed = [&](auto body, int reps, float *ms) -> int {
        for (int q = 0; q < 2; ++q)
            if (int r = body()) return r;
        VV_CHECK(hipEventRecord(e0, c->stream), "event");
        for (int q = 0; q < reps; ++q)
            if (int r = body()) return r;
        VV_CHECK(hipEventRecord(e1, c->stream), "event");
        VV_CHECK(hipEventSynchronize(e1), "event sync");
        VV_CHECK(hipEventElapsedTime(ms, e0, e1), "elapsed");
        return VV_OK;
    };
    int rc = VV_OK;
    do {
        // FFT rows: even R with R nx <= 8192
        if ((rc = launch_zero(c, c->X, c->N * c->B) == hipSuccess ? VV_OK : VV_ERR_HIP)) break;
        int bestR = 0;
        float bestT = 0.f;
        for (int R = 2; R * c->fft_n <= 8192; R <<= 1) {
            c->tune_fft_rows = R;
            float ms = 0.f;
            if ((rc = timed([&]() { return precond(c, c->X, c->D, nullptr); }, 5, &ms))) break;
            if (bestR == 0 || ms < bestT) {
                bestR = R;
                bestT = ms;
            }
        }
        if (rc) break;
        c->tune_fft_rows = bestR;
        // CGS vectors per thread, at basis size min(restart, 10)
        const int col = (c->p.restart < 10 ? c->p.restart : 10) - 1;
        if ((rc = launch_zero(c, c->basis, (size_t)(col + 2) * c->vs) == hipSuccess ? VV_OK : VV_ERR_HIP))
            break;
        int bestE = 0;
        bestT = 0.f;
        for (int E = 1; E <= 8; E <<= 1) {
            c->tune_cgs_ept = E;
            float ms = 0.f;
            auto body = [&]() -> int {
                VV_CHECK(launch_gmres_init(c, 1e-300, 0.0, 1e300, 1 << 30), "gmres init");
                VV_CHECK(launch_mdot(c, col), "mdot");
                VV_CHECK(launch_maxpy_norm(c, col), "maxpy");
                return VV_OK;
            };
            if ((rc = timed(body, 5, &ms))) break;
            if (bestE == 0 || ms < bestT) {
                bestE = E;
                bestT = ms;
            }
        }
        if (rc) break;
        c->tune_cgs_ept = bestE;
    } while (0);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->prof = prof;
    if (rc == VV_ERR_HIP && c->err.empty()) c->err = "autotune: kernel launch failed";
    if (rc) return rc;
    if (fft_rows) *fft_rows = c->tune_fft_rows;
    if (cgs_ept) *cgs_ept = c->tune_cgs_ept;
    return VV_OK;
}

int vv_set_hamiltonian(vv_ctx *c, const double *h0, const double *h1p, const double *h1h, const double *h2p,
                       const double *h2h)
{
    if (int r = check_ctx(c)) return r;
    if (!h0 || !h1p || !h1h || !h2p || !h2h) return fail(c, VV_ERR_INVALID, "null Hamiltonian factor");
    const int nxb = c->p.nx * c->B, nv = c->p.nv;
    const FactorCopy a{{h0, h1p, h1h, h2p, h2h}, {c->h0, c->h1p, c->h1h, c->h2p, c->h2h}, {nv, nxb, nxb, nxb, nxb}};
    const int n = nxb > nv ? nxb : nv;
    hipLaunchKernelGGL(factors_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, a);
    VV_CHECK(hipGetLastError(), "copy Hamiltonian factors");
    c->db1_valid = false;
    return VV_OK;
}

int vv_residual(vv_ctx *c, const double *f, const double *fh, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (y == f || y == fh) return fail(c, VV_ERR_INVALID, "residual: y must not alias f or fh");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, y), "residual");
    return diss_function(c, f, fh, y);
}

int vv_jacobian_apply(vv_ctx *c, const double *u, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (u == y) return fail(c, VV_ERR_INVALID, "jacobian_apply: u and y must not alias");
    VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, u, nullptr, y), "jacobian");
    return diss_jacobian(c, u, nullptr, y, nullptr);
}

int vv_bracket_add(vv_ctx *c, const double *x, double *y, double factor)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (x == y) return fail(c, VV_ERR_INVALID, "bracket_add: x and y must not alias");
    VV_CHECK(launch_bracket_add(c, x, y, factor), "bracket");
    return VV_OK;
}

int vv_precond_update(vv_ctx *c, const double *phi)
{
    if (int r = check_ctx(c)) return r;
    if (c->p.precond != VV_PRECOND_TENSOR_POTENTIAL) return VV_OK;
    if (!phi) return fail(c, VV_ERR_INVALID, "null phi");
    VV_CHECK(launch_potential_bands(c, phi), "potential bands");
    return VV_OK;
}

int vv_precond_apply(vv_ctx *c, const double *x, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    return precond(c, x, y, nullptr);
}

int vv_function(vv_ctx *c, const double *f, const double *fh, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y) return fail(c, VV_ERR_INVALID, "null field");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, c->X), "residual");
    if (int r = diss_function(c, f, fh, c->X)) return r;
    return precond(c, c->X, y, nullptr);
}

int vv_jacobian(vv_ctx *c, const double *u, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, u, nullptr, c->X), "jacobian");
    if (int r = diss_jacobian(c, u, nullptr, c->X, nullptr)) return r;
    return precond(c, c->X, y, nullptr);
}

int vv_function_norm(vv_ctx *c, const double *f, const double *fh, double *y, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y || !out) return fail(c, VV_ERR_INVALID, "null argument");
    if (y == f || y == fh) return fail(c, VV_ERR_INVALID, "function_norm: y must not alias f or fh");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, y), "residual");
    if (int r = diss_function(c, f, fh, y)) return r;
    VV_CHECK(launch_norms(c, y, ds_norms(c)), "norm");
    return read_scalars(c, ds_norms(c), (size_t)c->B, out);
}

int vv_batched_timestep(vv_ctx *c, const vv_batch_fields *s, const vv_ksp_params *kp, double snes_atol,
                        int32_t snes_max_iter, vv_instance_step *rec, double *history, int32_t history_rows,
                        int32_t *iterations)
{
    if (int r = check_ctx(c)) return r;
    if (!s || !kp || !rec || !s->f || !s->fh || !s->fl || !s->fy || !s->h0 || !s->h1 || !s->h1h || !s->h2 ||
        !s->h2h || !s->n || !s->phi || !s->phih)
        return fail(c, VV_ERR_INVALID, "null argument");
    if (snes_max_iter < 1 || history_rows < 0 || (history_rows > 0 && !history))
        return fail(c, VV_ERR_INVALID, "batched_timestep: bad iteration limit or history");
    const int B = c->B;
    const int64_t N = c->N, NB = c->N * c->B, XB = (int64_t)c->p.nx * c->B;
    std::vector<double> fo(B), po(B), pred(B);
    std::vector<int32_t> active(B, 1);
    std::vector<vv_ksp_result> res(B);
    auto residual = [&](std::vector<double> &out) -> int {
        if (int r = vv_residual_norms(c, s->f, s->fh, s->fy, fo.data(), po.data())) return r;
        for (int b = 0; b < B; ++b) out[b] = fo[b] + po[b];
        return VV_OK;
    };
    // make_history, update_previous and the residual before the loop
    VV_CHECK(launch_copy(c, s->f, s->fh, NB), "history");
    VV_CHECK(launch_copy(c, s->h1, s->h1h, XB), "history");
    VV_CHECK(launch_copy(c, s->h2, s->h2h, XB), "history");
    VV_CHECK(launch_copy(c, s->phi, s->phih, XB), "history");
    if (int r = vv_set_hamiltonian(c, s->h0, s->h1, s->h1h, s->h2, s->h2h)) return r;
    if (int r = residual(pred)) return r;
    for (int b = 0; b < B; ++b) rec[b] = vv_instance_step{0, 0, 0, pred[b], pred[b]};
    std::vector<double> now(B);
    int i = 0, rc = VV_OK;
    for (;;) {
        bool any = false;
        for (int b = 0; b < B; ++b) any = any || active[b];
        if (!any) break;
        ++i;
        if ((rc = vv_set_active(c, active.data()))) break;
        if (const hipError_t e = launch_copy(c, s->f, s->fl, NB)) {
            rc = hipfail(c, e, "backup");
            break;
        }
        if ((rc = vv_snes_ksponly(c, s->f, s->fh, kp, res.data()))) break;
        if ((rc = vv_moments(c, s->f, s->n, s->phi, s->h1, nullptr))) break;
        if ((rc = vv_set_hamiltonian(c, s->h0, s->h1, s->h1h, s->h2, s->h2h))) break;
        if ((rc = residual(now))) break;
        if (i <= history_rows)
            for (int b = 0; b < B; ++b) history[(size_t)(i - 1) * B + b] = active[b] ? now[b] : std::nan("");
        bool revert = false;
        for (int b = 0; b < B; ++b) {
            if (!active[b]) continue;
            rec[b].picard += 1;
            rec[b].gmres += res[b].its;
            const double prev = pred[b];
            pred[b] = now[b];
            if ((pred[b] > prev && i > 1) || pred[b] < snes_atol || i >= snes_max_iter) {
                active[b] = 0;
                if (pred[b] > prev) {
                    // the iterate of this instance is undone
                    if (const hipError_t e = launch_copy(c, s->fl + (size_t)b * N, s->f + (size_t)b * N, N)) {
                        rc = hipfail(c, e, "revert");
                        break;
                    }
                    rec[b].reverted = 1;
                    revert = true;
                }
            }
        }
        if (rc) break;
        if (revert) {
            if ((rc = vv_moments(c, s->f, s->n, s->phi, s->h1, nullptr))) break;
            if ((rc = vv_set_hamiltonian(c, s->h0, s->h1, s->h1h, s->h2, s->h2h))) break;
        }
    }
    const int r2 = vv_set_active(c, nullptr);
    if (rc) return rc;
    if (r2) return r2;
    for (int b = 0; b < B; ++b) rec[b].final = pred[b];
    if (iterations) *iterations = i;
    return VV_OK;
}

int vv_residual_norms(vv_ctx *c, const double *f, const double *fh, double *y, double *fnorm, double *pnorm)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y || !fnorm || !pnorm) return fail(c, VV_ERR_INVALID, "null argument");
    if (y == f || y == fh) return fail(c, VV_ERR_INVALID, "residual_norms: y must not alias f or fh");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, y), "residual");
    if (int r = diss_function(c, f, fh, y)) return r;
    VV_CHECK(launch_norms(c, y, ds_norms(c)), "norm");
    // ds_norms (B values) is followed by the Poisson residuals ds_pres (stride 16)
    const size_t B = (size_t)c->B;
    std::vector<double> tmp(B + 16 * B);
    if (int r = read_scalars(c, ds_norms(c), B + 16 * (B - 1) + 1, tmp.data())) return r;
    for (size_t b = 0; b < B; ++b) {
        fnorm[b] = tmp[b];
        pnorm[b] = tmp[B + 16 * b];
    }
    return VV_OK;
}

int vv_density(vv_ctx *c, const double *f, double *n)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !n) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_density(c, f, n), "density");
    return VV_OK;
}

int vv_velocity_moments(vv_ctx *c, const double *f, double *u, double *e)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !u || !e) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_velocity_moments(c, f, u, e), "velocity moments");
    return VV_OK;
}

// ---- explicit steppers of the initial guess ---------------------------------
namespace {
int explicit_setup(vv_ctx *c, double *x, int32_t niter)
{
    if (int r = check_ctx(c)) return r;
    if (!x) return fail(c, VV_ERR_INVALID, "null field");
    if (niter < 1) return fail(c, VV_ERR_INVALID, "niter must be >= 1");
    if (c->B != 1) return fail(c, VV_ERR_UNSUPPORTED, "the explicit steppers take a single-instance context");
    if (!c->ex) VV_CHECK(dalloc(c, &c->ex, 6 * (size_t)c->N), "alloc explicit-stepper scratch");
    return VV_OK;
}

// t = -{x, H} on rows 2..nv-3, 0 elsewhere
hipError_t tendency(vv_ctx *c, const double *x, const double *H, double *t)
{
    return launch_bracket_acc(c, x, nullptr, H, t, -1.0, true, nullptr);
}
}  // namespace

int vv_symplectic_kinetic(vv_ctx *c, double *x, const double *h0, double factor, int32_t niter)
{
    if (int r = explicit_setup(c, x, niter)) return r;
    if (!h0) return fail(c, VV_ERR_INVALID, "null h0");
    double *H = c->ex, *t = c->ex + c->N;
    VV_CHECK(launch_compose_h(c, h0, nullptr, nullptr, H), "compose H");
    VV_CHECK(tendency(c, x, H, t), "tendency");
    VV_CHECK(launch_stage(c, x, t, factor * c->p.ht / (double)niter, x), "update");
    return VV_OK;
}

int vv_symplectic_potential(vv_ctx *c, double *x, const double *h1, const double *h2, double factor, int32_t niter)
{
    if (int r = explicit_setup(c, x, niter)) return r;
    if (!h1 || !h2) return fail(c, VV_ERR_INVALID, "null h1 / h2");
    double *H = c->ex, *t = c->ex + c->N;
    VV_CHECK(launch_compose_h(c, nullptr, h1, h2, H), "compose H");
    VV_CHECK(tendency(c, x, H, t), "tendency");
    VV_CHECK(launch_stage(c, x, t, factor * c->p.ht / (double)niter, x), "update");
    return VV_OK;
}

int vv_rk4(vv_ctx *c, double *x, const double *h0, const double *h1, const double *h2, int32_t rule, int32_t niter)
{
    if (int r = explicit_setup(c, x, niter)) return r;
    if (!h0 || !h1 || !h2) return fail(c, VV_ERR_INVALID, "null Hamiltonian part");
    if (rule != VV_RK4_16 && rule != VV_RK4_18) return fail(c, VV_ERR_INVALID, "rule must be VV_RK4_16 or VV_RK4_18");
    const size_t N = (size_t)c->N;
    double *H = c->ex, *t1 = H + N, *t2 = t1 + N, *t3 = t2 + N, *t4 = t3 + N, *st = t4 + N;
    const double ht = c->p.ht, ni = (double)niter, s = ht / ni;
    VV_CHECK(launch_compose_h(c, h0, h1, h2, H), "compose H");
    VV_CHECK(tendency(c, x, H, t1), "tendency");
    if (rule == VV_RK4_16) {
        // PETScArakawaRungeKutta.pyx:41-57
        VV_CHECK(launch_stage(c, x, t1, 0.5 * ht / ni, st), "stage");
        VV_CHECK(tendency(c, st, H, t2), "tendency");
        VV_CHECK(launch_stage(c, x, t2, 0.5 * ht / ni, st), "stage");
        VV_CHECK(tendency(c, st, H, t3), "tendency");
        VV_CHECK(launch_stage(c, x, t3, 1.0 * ht / ni, st), "stage");
        VV_CHECK(tendency(c, st, H, t4), "tendency");
        VV_CHECK(launch_rk4_end(c, false, x, t1, t2, t3, t4, s), "update");
    } else {
        // PETScArakawaRungeKutta.pyx:82-98
        const double c13 = 1. / 3. * ht / ni;
        VV_CHECK(launch_stage(c, x, t1, c13, st), "stage");
        VV_CHECK(tendency(c, st, H, t2), "tendency");
        VV_CHECK(launch_rk18_stage3(c, x, t1, t2, c13, s, st), "stage");
        VV_CHECK(tendency(c, st, H, t3), "tendency");
        VV_CHECK(launch_rk18_stage4(c, x, t1, t2, t3, s, st), "stage");
        VV_CHECK(tendency(c, st, H, t4), "tendency");
        VV_CHECK(launch_rk4_end(c, true, x, t1, t2, t3, t4, s), "update");
    }
    return VV_OK;
}

int vv_set_dissipation(vv_ctx *c, const vv_dissipation *d)
{
    if (int r = check_ctx(c)) return r;
    if (!d) return fail(c, VV_ERR_INVALID, "null argument");
    if (d->kind < VV_DISSIPATION_NONE || d->kind > VV_DISSIPATION_FHH)
        return fail(c, VV_ERR_UNSUPPORTED, "dissipation must be none, collt, colle, ffh or fhh");
    // config.py:148-163: no dissipation object when coll_freq == 0
    const int kind = d->coll_freq == 0. ? VV_DISSIPATION_NONE : d->kind;
    if (kind != VV_DISSIPATION_NONE && c->B != 1)
        return fail(c, VV_ERR_UNSUPPORTED, "dissipation terms take a single-instance context");
    if (kind != VV_DISSIPATION_NONE && c->use_graph)
        return fail(c, VV_ERR_UNSUPPORTED, "switch graph replay off before setting a dissipation term");
    if ((kind == VV_DISSIPATION_FFH || kind == VV_DISSIPATION_FHH) && !c->fave) {
        VV_CHECK(dalloc(c, &c->fave, (size_t)c->N), "alloc double-bracket average");
        VV_CHECK(dalloc(c, &c->db1, (size_t)c->N), "alloc double-bracket scratch");
        VV_CHECK(dalloc(c, &c->db2, (size_t)c->N), "alloc double-bracket scratch");
        VV_CHECK(launch_zero(c, c->fave, c->N), "zero");    // the solver's Fave starts as a zero Vec
    }
    c->diss_kind = kind;
    c->coll_freq = d->coll_freq;
    c->coll_diff = d->coll_diff;
    c->coll_drag = d->coll_drag;
    c->db1_valid = false;
    return VV_OK;
}

int vv_set_moments(vv_ctx *c, const double *np, const double *up, const double *ep, const double *ap,
                   const double *nh, const double *uh, const double *eh, const double *ah)
{
    if (int r = check_ctx(c)) return r;
    const double *m[8] = {np, up, ep, ap, nh, uh, eh, ah};
    for (int k = 0; k < 8; ++k) c->mom[k] = m[k];
    return VV_OK;
}

int vv_collision_factor(vv_ctx *c, const double *n, const double *u, const double *e, double *a)
{
    if (int r = check_ctx(c)) return r;
    if (!n || !u || !e || !a) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_collision_factor(c, n, u, e, a), "collision factor");
    return VV_OK;
}

int vv_collision_add(vv_ctx *c, const double *f, double *y, const double *n, const double *u, const double *e,
                     const double *a, double factor)
{
    (void)e;
    if (int r = check_ctx(c)) return r;
    if (!f || !y || !n || !u || !a) return fail(c, VV_ERR_INVALID, "null argument");
    if (f == y) return fail(c, VV_ERR_INVALID, "collision_add: f and y must not alias");
    if (!diss_coll(c)) return fail(c, VV_ERR_UNSUPPORTED, "the context has no collision operator");
    VV_CHECK(launch_collision(c, f, nullptr, y, n, u, a, factor, nullptr), "collision");
    return VV_OK;
}

int vv_double_bracket_function(vv_ctx *c, const double *fave, double *y, double factor)
{
    if (int r = check_ctx(c)) return r;
    if (!fave || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (fave == y) return fail(c, VV_ERR_INVALID, "double_bracket_function: fave and y must not alias");
    if (c->diss_kind != VV_DISSIPATION_FFH && c->diss_kind != VV_DISSIPATION_FHH)
        return fail(c, VV_ERR_UNSUPPORTED, "the context has no double bracket");
    const double fac = -c->coll_freq * factor;
    VV_CHECK(launch_bracket_acc(c, fave, nullptr, nullptr, c->db2, 1.0, true, nullptr), "double bracket");
    if (c->diss_kind == VV_DISSIPATION_FFH)
        VV_CHECK(launch_bracket_acc(c, fave, nullptr, c->db2, y, fac, false, nullptr), "double bracket");
    else
        VV_CHECK(launch_bracket_acc(c, c->db2, nullptr, nullptr, y, fac, false, nullptr), "double bracket");
    return VV_OK;
}

int vv_double_bracket_jacobian(vv_ctx *c, const double *u, const double *fave, double *y, double factor)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !fave || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (u == y || fave == y) return fail(c, VV_ERR_INVALID, "double_bracket_jacobian: y must not alias u or fave");
    if (c->diss_kind != VV_DISSIPATION_FFH && c->diss_kind != VV_DISSIPATION_FHH)
        return fail(c, VV_ERR_UNSUPPORTED, "the context has no double bracket");
    const double fac = -c->coll_freq * factor;
    if (c->diss_kind == VV_DISSIPATION_FFH) {
        VV_CHECK(launch_bracket_acc(c, fave, nullptr, nullptr, c->db2, 1.0, true, nullptr), "double bracket");
        VV_CHECK(launch_bracket_acc(c, u, nullptr, c->db2, y, fac, false, nullptr), "double bracket");
        VV_CHECK(launch_bracket_acc(c, u, nullptr, nullptr, c->db2, 1.0, true, nullptr), "double bracket");
        VV_CHECK(launch_bracket_acc(c, fave, nullptr, c->db2, y, fac, false, nullptr), "double bracket");
    } else {
        VV_CHECK(launch_bracket_acc(c, u, nullptr, nullptr, c->db2, 1.0, true, nullptr), "double bracket");
        VV_CHECK(launch_bracket_acc(c, c->db2, nullptr, nullptr, y, fac, false, nullptr), "double bracket");
    }
    return VV_OK;
}

int vv_poisson_solve(vv_ctx *c, const double *n, double *phi)
{
    if (int r = check_ctx(c)) return r;
    if (!n || !phi) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_poisson(c, n, nullptr, phi, nullptr, nullptr, 0), "poisson");
    return VV_OK;
}

int vv_poisson_matrix_mult(vv_ctx *c, const double *x, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_poisson_mult(c, x, nullptr, y), "poisson mult");
    return VV_OK;
}

int vv_poisson_function_mult(vv_ctx *c, const double *x, const double *n, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !n || !y) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_poisson_mult(c, x, n, y), "poisson function");
    return VV_OK;
}

int vv_potential_to_hamiltonian(vv_ctx *c, const double *phi, double *h1)
{
    if (int r = check_ctx(c)) return r;
    if (!phi || !h1) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_pot_to_ham(c, phi, h1), "potential_to_hamiltonian");
    return VV_OK;
}

int vv_moments(vv_ctx *c, const double *f, double *n, double *phi, double *h1, double *pnorm)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !n || !phi || !h1) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_density_partial(c, f, nullptr), "density");
    VV_CHECK(launch_poisson(c, nullptr, n, phi, h1, ds_pres(c), 1), "poisson");
    if (!pnorm) return VV_OK;
    std::vector<double> tmp(16 * (size_t)c->B);
    if (int r = read_scalars(c, ds_pres(c), 16 * (size_t)(c->B - 1) + 1, tmp.data())) return r;
    for (int b = 0; b < c->B; ++b) pnorm[b] = tmp[16 * (size_t)b];
    return VV_OK;
}

int vv_moments_noalloc(vv_ctx *c, const double *f, double *n, double *phi, double *h1, double *pnorm)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !n || !phi || !h1) return fail(c, VV_ERR_INVALID, "null argument");
    const int64_t X = (int64_t)c->p.nx * c->B;
    if (X > kNoallocMaxX)
        return fail(c, VV_ERR_INVALID, "vv_moments_noalloc: nx * batch = " + std::to_string(X) +
                                           " exceeds its static scratch (nx * batch <= 4096)");
    if (!c->static_scratch) return fail(c, VV_ERR_HIP, "vv_moments_noalloc: static scratch address unknown");
    double *npart = c->static_scratch;
    double *work = c->static_scratch + kDensityChunks * X;
    // every partial and work entry is written before it is read: nothing to clear
    VV_CHECK(launch_density_partial(c, f, npart), "density");
    VV_CHECK(launch_poisson(c, nullptr, n, phi, h1, pnorm, 1, npart, work, 1), "poisson");
    return VV_OK;
}

int vv_norm2(vv_ctx *c, const double *x, int64_t len, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !out || len < 0) return fail(c, VV_ERR_INVALID, "bad argument");
    VV_CHECK(launch_norm2(c, x, len, ds_norms(c)), "norm");
    return read_scalars(c, ds_norms(c), 1, out);
}

int vv_dot(vv_ctx *c, const double *x, const double *y, int64_t len, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y || !out || len < 0) return fail(c, VV_ERR_INVALID, "bad argument");
    VV_CHECK(launch_dot(c, x, y, len, ds_norms(c)), "dot");
    return read_scalars(c, ds_norms(c), 1, out);
}

int vv_axpy(vv_ctx *c, double alpha, const double *x, double *y, int64_t len)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y || len < 0) return fail(c, VV_ERR_INVALID, "bad argument");
    VV_CHECK(launch_axpy(c, alpha, x, y, len), "axpy");
    return VV_OK;
}

int vv_ksp_solve(vv_ctx *c, const double *b, double *x, const vv_ksp_params *kp, vv_ksp_result *res)
{
    if (int r = check_ctx(c)) return r;
    if (!b || !x || !kp) return fail(c, VV_ERR_INVALID, "null argument");
    if (kp->max_it < 1) return fail(c, VV_ERR_INVALID, "max_it must be >= 1");
    return ksp_solve_impl(c, b, x, kp, res);
}

int vv_snes_ksponly(vv_ctx *c, double *f, const double *fh, const vv_ksp_params *kp, vv_ksp_result *res)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !kp) return fail(c, VV_ERR_INVALID, "null argument");
    if (kp->max_it < 1) return fail(c, VV_ERR_INVALID, "max_it must be >= 1");
    return snes_impl(c, f, fh, kp, res);
}

int vv_ksp_solve_mixed(vv_ctx *c, const double *b, double *x, const vv_ksp_params *kp, const vv_mixed_params *mp,
                       vv_ksp_result *res, int32_t *cycles)
{
    if (int r = check_ctx(c)) return r;
    if (!b || !x || !kp || !mp) return fail(c, VV_ERR_INVALID, "null argument");
    if (kp->max_it < 1) return fail(c, VV_ERR_INVALID, "max_it must be >= 1");
    if (!(mp->inner_rtol > 0.0 && mp->inner_rtol < 1.0) || mp->max_cycles < 1)
        return fail(c, VV_ERR_INVALID, "inner_rtol must be in (0, 1) and max_cycles >= 1");
    return ksp_solve_mixed_impl(c, b, x, kp, mp, res, cycles);
}

int vv_snes_ksponly_mixed(vv_ctx *c, double *f, const double *fh, const vv_ksp_params *kp, const vv_mixed_params *mp,
                          vv_ksp_result *res, int32_t *cycles)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !kp || !mp) return fail(c, VV_ERR_INVALID, "null argument");
    if (kp->max_it < 1) return fail(c, VV_ERR_INVALID, "max_it must be >= 1");
    if (!(mp->inner_rtol > 0.0 && mp->inner_rtol < 1.0) || mp->max_cycles < 1)
        return fail(c, VV_ERR_INVALID, "inner_rtol must be in (0, 1) and max_cycles >= 1");
    if (c->B != 1) return fail(c, VV_ERR_UNSUPPORTED, "the mixed-precision solve takes one instance");
    if (int r = vv_function(c, f, fh, c->rhs)) return r;
    vv_ksp_result kr{};
    if (int r = ksp_solve_mixed_impl(c, c->rhs, c->D, kp, mp, &kr, cycles)) return r;
    // SNESSolve_KSPONLY takes the step unless the linear solve diverged
    if (kr.reason > 0) VV_CHECK(launch_axpy(c, -1.0, c->D, f, c->N), "update");
    if (res) *res = kr;
    return VV_OK;
}

int vv_jacobian_apply32(vv_ctx *c, const float *u, float *y)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (u == y) return fail(c, VV_ERR_INVALID, "jacobian_apply32: u and y must not alias");
    if (int r = no_diss(c, "the FP32 Jacobian")) return r;
    VV_PLAUNCH(P_JAC32, 8.0 * (double)c->N * c->B, -1, launch_jacobian32(c, u, nullptr, y), "jacobian32");
    return VV_OK;
}

int vv_precond_apply32(vv_ctx *c, const float *x, float *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (c->p.precond == VV_PRECOND_NONE)
        return fail(c, VV_ERR_UNSUPPORTED, "precond_apply32 needs a tensor-product preconditioner");
    return precond32(c, x, y, nullptr);
}

int vv_precond_apply_t(vv_ctx *c, const double *x, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (c->p.precond == VV_PRECOND_NONE)
        return fail(c, VV_ERR_UNSUPPORTED, "the transposed preconditioner needs a tensor-product preconditioner");
    return precond_t(c, x, y, nullptr);
}

int vv_linear_apply(vv_ctx *c, const double *u, const double *fbar, double *y, int32_t which)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (u == y || (fbar && (fbar == y))) return fail(c, VV_ERR_INVALID, "linear_apply: y must not alias u or fbar");
    if (int r = no_diss(c, "the linearised operator")) return r;
    const double it = c->ht_inv;
    switch (which) {
    case VV_LIN_A: return linear_apply(c, u, nullptr, fbar, y, it, 0.5, 0.5, false);
    case VV_LIN_A0: return linear_apply(c, u, nullptr, fbar, y, -it, 0.5, 0.5, false);
    case VV_LIN_AT: return linear_apply(c, u, nullptr, fbar, y, it, -0.5, 0.5, true);
    case VV_LIN_A0T: return linear_apply(c, u, nullptr, fbar, y, -it, -0.5, 0.5, true);
    default: return fail(c, VV_ERR_INVALID, "which must be VV_LIN_A, _A0, _AT or _A0T");
    }
}

int vv_adjoint_solve(vv_ctx *c, const double *lam, const double *fbar, double *mu, const vv_ksp_params *kp,
                     vv_ksp_result *res)
{
    if (int r = check_ctx(c)) return r;
    if (!lam || !mu || !kp) return fail(c, VV_ERR_INVALID, "null argument");
    if (kp->max_it < 1) return fail(c, VV_ERR_INVALID, "max_it must be >= 1");
    if (c->p.precond == VV_PRECOND_NONE)
        return fail(c, VV_ERR_UNSUPPORTED, "the adjoint solve needs a tensor-product preconditioner");
    if (c->use_graph) return fail(c, VV_ERR_UNSUPPORTED, "switch graph replay off for the adjoint solve");
    if (int r = no_diss(c, "the adjoint solve")) return r;
    if (int r = ensure_linear(c)) return r;
    if (int r = precond_t(c, lam, c->rhs, nullptr)) return r;
    c->op_mode = 1;
    c->op_fbar = fbar;
    const int r = ksp_solve_impl(c, c->rhs, mu, kp, res);
    c->op_mode = 0;
    c->op_fbar = nullptr;
    return r;
}

int vv_adjoint_step(vv_ctx *c, const double *lam1, const double *fbar, double *lam0, double *mu,
                    const vv_ksp_params *kp, vv_ksp_result *res)
{
    if (int r = check_ctx(c)) return r;
    if (!lam1 || !lam0 || !mu || !kp) return fail(c, VV_ERR_INVALID, "null argument");
    if (lam0 == lam1 || lam0 == mu || mu == lam1 || (fbar && (fbar == lam0 || fbar == mu)))
        return fail(c, VV_ERR_INVALID, "adjoint_step: lam1, lam0, mu and fbar must be distinct");
    if (int r = vv_adjoint_solve(c, lam1, fbar, mu, kp, res)) return r;
    // lam0 = -A0^T mu = mu/dt + 1/2 {mu~, H} - 1/2 h1[{mu~, f_bar}]
    return linear_apply(c, mu, nullptr, fbar, lam0, c->ht_inv, 0.5, -0.5, true);
}

int vv_invariants(vv_ctx *c, const double *f, const double *h0, const double *h1, const double *h2,
                  const double *phi, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !h0 || !h1 || !h2 || !phi || !out) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_invariants(c, f, h0, h1, h2, phi, ds_inv(c)), "invariants");
    return read_scalars(c, ds_inv(c), 6 * (size_t)c->B, out);
}

int vv_profile_enable(vv_ctx *c, int enable)
{
    if (int r = check_ctx(c)) return r;
    c->prof = enable;
    return VV_OK;
}

int vv_profile_sample(vv_ctx *c, int every)
{
    if (int r = check_ctx(c)) return r;
    if (every < 1) return fail(c, VV_ERR_INVALID, "every must be >= 1");
    Prof *p = prof_of(c);
    p->every = every;
    for (auto &n : p->seen) n = 0;
    return VV_OK;
}

int vv_profile_read(vv_ctx *c, const char *kernel, double *ms, int64_t *launches, double *bytes)
{
    if (int r = check_ctx(c)) return r;
    if (!kernel || !ms || !launches || !bytes) return fail(c, VV_ERR_INVALID, "null argument");
    int kind = -1;
    for (int k = 0; k < P_NKINDS; ++k)
        if (std::strcmp(kernel, kProfNames[k]) == 0) kind = k;
    if (kind < 0) return fail(c, VV_ERR_INVALID, std::string("unknown kernel ") + kernel);
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    Prof *p = prof_of(c);
    double tot = 0.0, by = 0.0;
    int64_t cnt = 0;
    std::vector<ProfRec> keep;
    for (auto &r : p->recs) {
        if (r.kind == kind) {
            float t = 0.f;
            // launches that returned at once (cycle already finished) are not counted
            if (r.bytes > 0.0 && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
                tot += t;
                by += r.bytes;
                ++cnt;
            }
            p->pool.push_back(r.a);
            p->pool.push_back(r.b);
        } else {
            keep.push_back(r);
        }
    }
    p->recs.swap(keep);
    *ms = tot;
    *launches = cnt;
    *bytes = by;
    return VV_OK;
}

}  // extern "C"

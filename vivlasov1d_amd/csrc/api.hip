// C ABI of libvv1d.so: context set-up, the operator entry points and the
// native GMRES / SNES(ksponly) drivers.  See include/vv1d.h.
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
    P_JAC32, P_RFFT32, P_TRI32, P_IRFFT32, P_MDOT32, P_MAXPY32, P_NKINDS
};
const char *kProfNames[P_NKINDS] = {"jacobian",   "residual", "rfft",      "tridiag", "irfft",
                                    "mdot",       "maxpy_norm", "jacobian32", "rfft32",  "tridiag32",
                                    "irfft32",    "mdot32",   "maxpy_norm32"};

struct ProfRec {
    int kind;
    int tag;          // Arnoldi iteration within the cycle, -1 outside GMRES
    double bytes;     // algorithmic HBM bytes of this launch
    hipEvent_t a, b;
};

struct Prof {
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
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
    ProfScope(vv_ctx *cc, int k, double by, int tg = -1) : c(cc), kind(k), tag(tg), bytes(by)
    {
        if ((c->prof >> kind) & 1) {
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

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename T>
hipError_t dalloc(T **p, size_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T) > 0 ? count * sizeof(T) : 16);
}

int read_scalar(vv_ctx *c, const double *dev, double *out)
{
    VV_CHECK(hipMemcpyAsync(&c->host_status[1].res, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream),
             "copy scalar");
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    *out = c->host_status[1].res;
    return VV_OK;
}

int check_ctx(vv_ctx *c)
{
    if (!c) return fail(nullptr, VV_ERR_INVALID, "null context");
    return VV_OK;
}

// algorithmic HBM bytes per launch (compulsory reads + writes of each kernel)
double b_field(vv_ctx *c) { return 8.0 * (double)c->N; }
double b_spec(vv_ctx *c) { return 16.0 * (double)c->nmodes * (double)c->p.nv; }

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
int precond(vv_ctx *c, const double *x, double *y, const int32_t *done)
{
    if (c->p.precond == VV_PRECOND_NONE) {
        if (x != y) VV_CHECK(launch_copy(c, x, y, c->N), "copy");
        return VV_OK;
    }
    VV_PLAUNCH(P_RFFT, b_field(c) + b_spec(c), -1, launch_rfft(c, x, c->spec, done), "rfft");
    VV_PLAUNCH(P_TRI, 2 * b_spec(c), -1, launch_tridiag_k(c, c->spec, done), "tridiag");
    VV_PLAUNCH(P_IRFFT, b_field(c) + b_spec(c), -1, launch_irfft(c, c->spec, y, done), "irfft");
    return VV_OK;
}

// one Arnoldi step of GMRES on P^-1 J: V_{it+1} from V_it
int arnoldi_step(vv_ctx *c, int it)
{
    const int32_t *done = &c->gs->st.done;
    double *v = c->basis + (size_t)it * c->N;
    double *w = c->basis + (size_t)(it + 1) * c->N;
    const double F = b_field(c), S = b_spec(c);
    VV_PLAUNCH(P_JAC, 2 * F, it, launch_jacobian(c, v, &c->gs->scale[it], c->X), "jacobian");
    VV_PLAUNCH(P_RFFT, F + S, it, launch_rfft(c, c->X, c->spec, done), "rfft");
    VV_PLAUNCH(P_TRI, 2 * S, it, launch_tridiag_k(c, c->spec, done), "tridiag");
    VV_PLAUNCH(P_IRFFT, S + F, it, launch_irfft(c, c->spec, w, done), "irfft");
    VV_PLAUNCH(P_MDOT, F * (it + 2), it, launch_mdot(c, it), "mdot");
    VV_PLAUNCH(P_MAXPY, F * (it + 3), it, launch_maxpy_norm(c, it), "maxpy_norm");
    return VV_OK;
}

// y = P^-1 x in FP32 (complex64 spectrum in c->spec)
int precond32(vv_ctx *c, const float *x, float *y, const int32_t *done)
{
    float2 *spec = reinterpret_cast<float2 *>(c->spec);
    const double F = 4.0 * (double)c->N, S = 0.5 * b_spec(c);
    VV_PLAUNCH(P_RFFT32, F + S, -1, launch_rfft32(c, x, spec, done), "rfft32");
    VV_PLAUNCH(P_TRI32, 2 * S, -1, launch_tridiag32(c, spec, done), "tridiag32");
    VV_PLAUNCH(P_IRFFT32, S + F, -1, launch_irfft32(c, spec, y, done), "irfft32");
    return VV_OK;
}

// one Arnoldi step of the FP32 cycle: V_{it+1} = P32^-1 J32 v_it, CGS in FP64
// arithmetic on the FP32 basis
int arnoldi_step32(vv_ctx *c, int it)
{
    const int32_t *done = &c->gs->st.done;
    float *v = c->basis32 + (size_t)it * c->N;
    float *w = c->basis32 + (size_t)(it + 1) * c->N;
    float *X = reinterpret_cast<float *>(c->X);
    float2 *spec = reinterpret_cast<float2 *>(c->spec);
    const double F = 4.0 * (double)c->N, S = 0.5 * b_spec(c);
    VV_PLAUNCH(P_JAC32, 2 * F, it, launch_jacobian32(c, v, &c->gs->scale[it], X), "jacobian32");
    VV_PLAUNCH(P_RFFT32, F + S, it, launch_rfft32(c, X, spec, done), "rfft32");
    VV_PLAUNCH(P_TRI32, 2 * S, it, launch_tridiag32(c, spec, done), "tridiag32");
    VV_PLAUNCH(P_IRFFT32, S + F, it, launch_irfft32(c, spec, w, done), "irfft32");
    VV_PLAUNCH(P_MDOT32, F * (it + 2), it, launch_mdot32(c, it), "mdot32");
    VV_PLAUNCH(P_MAXPY32, F * (it + 3), it, launch_maxpy_norm32(c, it), "maxpy_norm32");
    return VV_OK;
}

// Arnoldi steps of one GMRES cycle, polling the device status every two
// steps (the status copy of step k is read while steps k+1, k+2 run); returns
// once the cycle has finished on the device, with its status in
// host_status[0].  Steps enqueued past the end of the cycle return at once on
// the device.
int run_cycle(vv_ctx *c, bool fp32)
{
    const int max_k = c->p.restart;
    const int chunk = 2;
    const size_t prof_from = c->prof ? prof_of(c)->recs.size() : 0;
    int it = 0, pend = -1, slot = 0;
    bool stop = false;
    while (it < max_k && !stop) {
        const int end = it + chunk < max_k ? it + chunk : max_k;
        for (; it < end; ++it) {
            int r = fp32 ? arnoldi_step32(c, it) : arnoldi_step(c, it);
            if (r) return r;
        }
        VV_CHECK(hipMemcpyAsync(&c->host_status[slot], &c->gs->st, sizeof(GmresStatus), hipMemcpyDeviceToHost,
                                c->stream), "status copy");
        VV_CHECK(hipEventRecord(c->ev[slot], c->stream), "event");
        if (pend >= 0) {
            VV_CHECK(hipEventSynchronize(c->ev[pend]), "event sync");
            if (c->host_status[pend].done) stop = true;
        }
        pend = slot;
        slot ^= 1;
    }
    VV_CHECK(hipMemcpyAsync(&c->host_status[0], &c->gs->st, sizeof(GmresStatus), hipMemcpyDeviceToHost, c->stream),
             "status copy");
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    prof_mark_idle(c, prof_from, c->host_status[0].it);
    return VV_OK;
}

int ksp_solve_impl(vv_ctx *c, const double *b, double *x, const vv_ksp_params *kp, vv_ksp_result *res)
{
    if (c->p.precond != VV_PRECOND_TENSOR_KINETIC)
        return fail(c, VV_ERR_UNSUPPORTED, "vv_ksp_solve needs the tensor_kinetic preconditioner");
    const int64_t N = c->N;
    VV_CHECK(hipMemsetAsync(x, 0, N * sizeof(double), c->stream), "zero x");
    VV_CHECK(launch_gmres_init(c, kp->rtol, kp->atol, kp->dtol, kp->max_it), "gmres init");
    VV_CHECK(launch_copy(c, b, c->basis, N), "copy b");
    bool first = true;
    GmresStatus st{};
    for (;;) {
        if (!first) {
            // KSPInitialResidual after a restart: r = b - (P^-1 J) x
            VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, x, nullptr, c->X), "jacobian");
            int r = precond(c, c->X, c->basis, nullptr);
            if (r) return r;
            VV_CHECK(launch_residual_restart(c, b, c->basis), "restart residual");
        }
        first = false;
        VV_CHECK(launch_norm_start(c), "cycle start");
        if (int r = run_cycle(c, false)) return r;
        VV_CHECK(launch_build_solution(c, x), "build solution");
        VV_CHECK(hipMemcpyAsync(&c->host_status[0], &c->gs->st, sizeof(GmresStatus), hipMemcpyDeviceToHost,
                                c->stream), "status copy");
        VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
        st = c->host_status[0];
        if (st.reason != 0) break;
        if (st.its >= kp->max_it) {
            st.reason = VV_KSP_DIVERGED_ITS;
            break;
        }
    }
    if (res) {
        res->its = st.its;
        res->reason = st.reason;
        res->rnorm = st.res;
        res->rnorm0 = st.rnorm0;
    }
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
    if (c->p.precond != VV_PRECOND_TENSOR_KINETIC)
        return fail(c, VV_ERR_UNSUPPORTED, "vv_ksp_solve_mixed needs the tensor_kinetic preconditioner");
    const int64_t N = c->N;
    VV_CHECK(hipMemsetAsync(x, 0, N * sizeof(double), c->stream), "zero x");
    VV_CHECK(launch_norm2(c, b, N, c->dscalar), "norm");
    double rnorm0 = 0.0;
    if (int r = read_scalar(c, c->dscalar, &rnorm0)) return r;
    const double ttol = std::fmax(kp->rtol * rnorm0, kp->atol);
    double rnorm = rnorm0;
    const double *rsrc = b;
    int its = 0, rounds = 0, reason = 0;
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
        if (int r = run_cycle(c, true)) return r;
        its += c->host_status[0].it;
        ++rounds;
        VV_CHECK(launch_build_solution32(c, x), "build solution");
        // true FP64 residual r = b - P^-1 J x
        VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, x, nullptr, c->X), "jacobian");
        if (int r = precond(c, c->X, c->R, nullptr)) return r;
        VV_CHECK(launch_residual_restart(c, b, c->R), "residual");
        VV_CHECK(launch_norm2(c, c->R, N, c->dscalar), "norm");
        if (int r = read_scalar(c, c->dscalar, &rnorm)) return r;
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

}  // namespace

extern "C" {

const char *vv_last_error(const vv_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

int vv_ctx_create(const vv_params *p, void *stream, vv_ctx **out)
{
    vv_ctx *c = nullptr;
    if (!p || !out) return fail(nullptr, VV_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!is_pow2(p->nx) || p->nx < 8 || p->nx > kMaxNx)
        return fail(nullptr, VV_ERR_INVALID, "nx must be a power of two in [8, 4096]");
    if (p->nv < 8 || p->nv > 4096) return fail(nullptr, VV_ERR_INVALID, "nv must be in [8, 4096]");
    if (p->bracket != VV_BRACKET_J1 && p->bracket != VV_BRACKET_J2 && p->bracket != VV_BRACKET_J4)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "bracket must be arakawaj1, arakawaj2 or arakawaj4");
    if (p->precond != VV_PRECOND_NONE && p->precond != VV_PRECOND_TENSOR_KINETIC)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "preconditioner must be none or tensor_kinetic");
    if (p->laplace != VV_LAPLACE_CFD2 && p->laplace != VV_LAPLACE_CFD4)
        return fail(nullptr, VV_ERR_UNSUPPORTED, "laplace operator must be cfd2 or cfd4");
    if (p->restart < 1 || p->restart > kMaxRestart) return fail(nullptr, VV_ERR_INVALID, "restart must be in [1, 64]");
    if (!(p->hx > 0) || !(p->hv > 0) || !(p->ht > 0)) return fail(nullptr, VV_ERR_INVALID, "hx, hv, ht must be > 0");
    if (!p->v) return fail(nullptr, VV_ERR_INVALID, "v grid missing");
    if (p->precond == VV_PRECOND_TENSOR_KINETIC && !p->eigen)
        return fail(nullptr, VV_ERR_INVALID, "kinetic eigenvalues missing");

    c = new (std::nothrow) vv_ctx();
    if (!c) return fail(nullptr, VV_ERR_INVALID, "out of host memory");
    c->p = *p;
    c->p.v = nullptr;
    c->p.eigen = nullptr;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    const int nx = p->nx, nv = p->nv;
    c->N = (int64_t)nx * nv;
    c->nmodes = nx / 2 + 1;
    c->logn = 0;
    while ((1 << c->logn) < nx) ++c->logn;
    // Grid.create (Grid.pyx:40-48) and the bracket / band factors
    const double hx = p->hx, hv = p->hv;
    c->ht_inv = 1. / p->ht;
    c->hx_inv = 1. / hx;
    c->hv_inv = 1. / hv;
    const double hx2 = hx * hx, hv2 = hv * hv;
    c->hx2_inv = 1. / hx2;
    c->af1 = c->hx_inv * c->hv_inv / 12.;          // PoissonBracket.pyx:88
    c->af2 = c->hx_inv * c->hv_inv / 24.;          // PoissonBracket.pyx:209
    c->afpc = 0.5 * c->hx_inv * c->hv_inv / 12.;   // KineticFast.pyx:251
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

    const int nfb = fft_blocks(c), ncb = cgs_blocks(c->N);
    int nrows = (nfb > kReduceBlocks ? nfb : kReduceBlocks);
    if (ncb > nrows) nrows = ncb;
    VV_CCHECK(dalloc(&c->h0, nv), "alloc h0");
    VV_CCHECK(dalloc(&c->h1p, nx), "alloc h1p");
    VV_CCHECK(dalloc(&c->h1h, nx), "alloc h1h");
    VV_CCHECK(dalloc(&c->h2p, nx), "alloc h2p");
    VV_CCHECK(dalloc(&c->h2h, nx), "alloc h2h");
    VV_CCHECK(dalloc(&c->eigen, c->nmodes), "alloc eigen");
    VV_CCHECK(dalloc(&c->gm, nv), "alloc gm");
    VV_CCHECK(dalloc(&c->g0, nv), "alloc g0");
    VV_CCHECK(dalloc(&c->gp, nv), "alloc gp");
    VV_CCHECK(dalloc(&c->twiddle, nx), "alloc twiddle");
    VV_CCHECK(dalloc(&c->twiddle32, nx), "alloc twiddle32");
    VV_CCHECK(dalloc(&c->spec, (size_t)c->nmodes * nv), "alloc spectrum");
    VV_CCHECK(dalloc(&c->pinv, nx), "alloc pinv");
    VV_CCHECK(dalloc(&c->pwork, 3 * nx + 1), "alloc Poisson scratch");
    VV_CCHECK(dalloc(&c->X, c->N), "alloc X");
    VV_CCHECK(dalloc(&c->B, c->N), "alloc B");
    VV_CCHECK(dalloc(&c->D, c->N), "alloc D");
    VV_CCHECK(dalloc(&c->basis, (size_t)(p->restart + 1) * c->N), "alloc Krylov basis");
    // the mixed solver's FP32 basis and FP64 residual share this allocation:
    // (restart+1) N floats, then N doubles (fits for restart >= 1)
    c->basis32 = reinterpret_cast<float *>(c->basis);
    c->R = c->basis + (size_t)(p->restart + 1) * (size_t)(c->N / 2);
    VV_CCHECK(dalloc(&c->partial, (size_t)nrows * (kMaxRestart + 2)), "alloc partial");
    VV_CCHECK(dalloc(&c->dscalar, 16), "alloc scalars");
    VV_CCHECK(dalloc(&c->npart, (size_t)64 * nx), "alloc density partials");
    VV_CCHECK(dalloc(&c->gs, 1), "alloc gmres state");
    VV_CCHECK(hipHostMalloc(reinterpret_cast<void **>(&c->host_status), 2 * sizeof(GmresStatus)), "pinned status");
    VV_CCHECK(hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming), "event");
    VV_CCHECK(hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming), "event");
    VV_CCHECK(hipMemset(c->gs, 0, sizeof(GmresState)), "zero gmres state");
    VV_CCHECK(hipMemset(c->partial, 0, (size_t)nrows * (kMaxRestart + 2) * sizeof(double)), "zero partials");

    // host tables
    std::vector<double> hzero(nx > nv ? nx : nv, 0.0);
    std::vector<double> gm(nv), g0(nv), gp(nv);
    for (int j = 0; j < nv; ++j) {
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
    std::vector<double2> tw(nx);
    std::vector<float2> tw32(nx);
    for (int m = 0; m < nx; ++m) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)nx;
        tw[m].x = (double)cosl(ang);
        tw[m].y = (double)sinl(ang);
        tw32[m].x = (float)cosl(ang);
        tw32[m].y = (float)sinl(ang);
    }
    // mean-free pseudo-inverse of the periodic Laplacian: c_m = 1/nx sum_{k>0} cos(2 pi k m/nx) / lambda_k
    std::vector<long double> ct(nx), lam(nx);
    for (int q = 0; q < nx; ++q)
        ct[q] = cosl(2.0L * 3.141592653589793238462643383279502884L * (long double)q / (long double)nx);
    for (int k = 1; k < nx; ++k) {
        const long double th2 = 3.141592653589793238462643383279502884L * (long double)k / (long double)nx;
        const long double s1 = sinl(th2), s2 = sinl(2.0L * th2);
        const long double h2 = (long double)hx * (long double)hx;
        // CFD4 symbol (54 - 64 cos t + 10 cos 2t)/12 = (128 sin^2(t/2) - 20 sin^2 t)/12; CFD2: 4 sin^2(t/2)
        lam[k] = (p->laplace == VV_LAPLACE_CFD4) ? (128.0L * s1 * s1 - 20.0L * s2 * s2) / (12.0L * h2)
                                                 : (4.0L * s1 * s1) / h2;
    }
    std::vector<double> pinv(nx);
    for (int m = 0; m < nx; ++m) {
        long double s = 0.0L;
        for (int k = 1; k < nx; ++k) s += ct[(int)(((long long)k * m) % nx)] / lam[k];
        pinv[m] = (double)(s / (long double)nx);
    }
    VV_CCHECK(hipMemcpy(c->h0, hzero.data(), nv * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->h1p, hzero.data(), nx * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->h1h, hzero.data(), nx * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->h2p, hzero.data(), nx * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->h2h, hzero.data(), nx * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->gm, gm.data(), nv * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->g0, g0.data(), nv * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->gp, gp.data(), nv * sizeof(double), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->eigen, eig.data(), c->nmodes * sizeof(double2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->twiddle, tw.data(), nx * sizeof(double2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->twiddle32, tw32.data(), nx * sizeof(float2), hipMemcpyHostToDevice), "upload");
    VV_CCHECK(hipMemcpy(c->pinv, pinv.data(), nx * sizeof(double), hipMemcpyHostToDevice), "upload");
#undef VV_CCHECK
    *out = c;
    return VV_OK;
}

int vv_ctx_destroy(vv_ctx *c)
{
    if (!c) return VV_OK;
    (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->h0, c->h1p, c->h1h, c->h2p, c->h2h, c->eigen, c->gm, c->g0, c->gp, c->twiddle, c->twiddle32,
                    c->spec,
                    c->pinv, c->pwork, c->X, c->B, c->D, c->basis, c->partial, c->dscalar, c->npart, c->gs};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->host_status) (void)hipHostFree(c->host_status);
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
    if (!c) return 0;
    const int64_t nx = c->p.nx, nv = c->p.nv;
    return (int64_t)sizeof(double) * ((c->p.restart + 4) * c->N + 64 * nx + 4 * nx + 4 * nv) +
           (int64_t)sizeof(double2) * ((int64_t)c->nmodes * nv + nx + c->nmodes) + (int64_t)sizeof(float2) * nx +
           (int64_t)sizeof(GmresState);
}

int vv_set_hamiltonian(vv_ctx *c, const double *h0, const double *h1p, const double *h1h, const double *h2p,
                       const double *h2h)
{
    if (int r = check_ctx(c)) return r;
    if (!h0 || !h1p || !h1h || !h2p || !h2h) return fail(c, VV_ERR_INVALID, "null Hamiltonian factor");
    const size_t bx = c->p.nx * sizeof(double), bv = c->p.nv * sizeof(double);
    VV_CHECK(hipMemcpyAsync(c->h0, h0, bv, hipMemcpyDeviceToDevice, c->stream), "copy h0");
    VV_CHECK(hipMemcpyAsync(c->h1p, h1p, bx, hipMemcpyDeviceToDevice, c->stream), "copy h1p");
    VV_CHECK(hipMemcpyAsync(c->h1h, h1h, bx, hipMemcpyDeviceToDevice, c->stream), "copy h1h");
    VV_CHECK(hipMemcpyAsync(c->h2p, h2p, bx, hipMemcpyDeviceToDevice, c->stream), "copy h2p");
    VV_CHECK(hipMemcpyAsync(c->h2h, h2h, bx, hipMemcpyDeviceToDevice, c->stream), "copy h2h");
    return VV_OK;
}

int vv_residual(vv_ctx *c, const double *f, const double *fh, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y) return fail(c, VV_ERR_INVALID, "null field");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, y), "residual");
    return VV_OK;
}

int vv_jacobian_apply(vv_ctx *c, const double *u, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, u, nullptr, y), "jacobian");
    return VV_OK;
}

int vv_bracket_add(vv_ctx *c, const double *x, double *y, double factor)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (x == y) return fail(c, VV_ERR_INVALID, "bracket_add: x and y must not alias");
    VV_CHECK(launch_bracket_add(c, x, y, factor), "bracket");
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
    return precond(c, c->X, y, nullptr);
}

int vv_jacobian(vv_ctx *c, const double *u, double *y)
{
    if (int r = check_ctx(c)) return r;
    if (!u || !y) return fail(c, VV_ERR_INVALID, "null field");
    VV_PLAUNCH(P_JAC, 2 * b_field(c), -1, launch_jacobian(c, u, nullptr, c->X), "jacobian");
    return precond(c, c->X, y, nullptr);
}

int vv_function_norm(vv_ctx *c, const double *f, const double *fh, double *y, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !fh || !y || !out) return fail(c, VV_ERR_INVALID, "null argument");
    VV_PLAUNCH(P_RES, 3 * b_field(c), -1, launch_residual(c, f, fh, y), "residual");
    VV_CHECK(launch_norm2(c, y, c->N, c->dscalar), "norm");
    return read_scalar(c, c->dscalar, out);
}

int vv_density(vv_ctx *c, const double *f, double *n)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !n) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_density(c, f, n), "density");
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
    VV_CHECK(launch_density_partial(c, f), "density");
    VV_CHECK(launch_poisson(c, nullptr, n, phi, h1, c->dscalar + 1, 1), "poisson");
    if (pnorm) return read_scalar(c, c->dscalar + 1, pnorm);
    return VV_OK;
}

int vv_norm2(vv_ctx *c, const double *x, int64_t len, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !out || len < 0) return fail(c, VV_ERR_INVALID, "bad argument");
    VV_CHECK(launch_norm2(c, x, len, c->dscalar), "norm");
    return read_scalar(c, c->dscalar, out);
}

int vv_dot(vv_ctx *c, const double *x, const double *y, int64_t len, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y || !out || len < 0) return fail(c, VV_ERR_INVALID, "bad argument");
    VV_CHECK(launch_dot(c, x, y, len, c->dscalar), "dot");
    return read_scalar(c, c->dscalar, out);
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
    if (int r = vv_function(c, f, fh, c->B)) return r;
    if (int r = ksp_solve_impl(c, c->B, c->D, kp, res)) return r;
    VV_CHECK(launch_axpy_cond(c, -1.0, c->D, f, c->N), "update");
    return VV_OK;
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
    if (int r = vv_function(c, f, fh, c->B)) return r;
    vv_ksp_result kr{};
    if (int r = ksp_solve_mixed_impl(c, c->B, c->D, kp, mp, &kr, cycles)) return r;
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
    VV_PLAUNCH(P_JAC32, 8.0 * (double)c->N, -1, launch_jacobian32(c, u, nullptr, y), "jacobian32");
    return VV_OK;
}

int vv_precond_apply32(vv_ctx *c, const float *x, float *y)
{
    if (int r = check_ctx(c)) return r;
    if (!x || !y) return fail(c, VV_ERR_INVALID, "null field");
    if (c->p.precond != VV_PRECOND_TENSOR_KINETIC)
        return fail(c, VV_ERR_UNSUPPORTED, "precond_apply32 needs the tensor_kinetic preconditioner");
    return precond32(c, x, y, nullptr);
}

int vv_invariants(vv_ctx *c, const double *f, const double *h0, const double *h1, const double *h2,
                  const double *phi, double *out)
{
    if (int r = check_ctx(c)) return r;
    if (!f || !h0 || !h1 || !h2 || !phi || !out) return fail(c, VV_ERR_INVALID, "null argument");
    VV_CHECK(launch_invariants(c, f, h0, h1, h2, phi, c->dscalar + 2), "invariants");
    VV_CHECK(hipStreamSynchronize(c->stream), "synchronize");
    double tmp[6];
    VV_CHECK(hipMemcpy(tmp, c->dscalar + 2, 6 * sizeof(double), hipMemcpyDeviceToHost), "copy");
    for (int q = 0; q < 6; ++q) out[q] = tmp[q];
    return VV_OK;
}

int vv_profile_enable(vv_ctx *c, int enable)
{
    if (int r = check_ctx(c)) return r;
    c->prof = enable;
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

// axs_api.hip -- the extern "C" entry points of include/axs_gpu.h: context
// lifecycle, state transfer, the step loop and the sampling calls.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <thread>

#include "axs_internal.hpp"

namespace axs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// coefficient table of build_coefficient_matrix (s_integrate.cpp:9-14), in
// float as there, widened to double as the reference's data_t initialiser
int stencil_coefficients(int stencil, int ndims, double *c)
{
    float nd = (float)ndims;
    switch (stencil) {
    case 0:
        c[0] = (double)(-2.0f * nd);
        c[1] = (double)1.0f;
        return 2;
    case 1:
        c[0] = (double)(-5.0f / 2.0f * nd);
        c[1] = (double)(4.0f / 3.0f);
        c[2] = (double)(-1.0f / 12.0f);
        return 3;
    case 2:
        c[0] = (double)(-49.0f / 18.0f * nd);
        c[1] = (double)(3.0f / 2.0f);
        c[2] = (double)(-3.0f / 20.0f);
        c[3] = (double)(1.0f / 90.0f);
        return 4;
    case 3:
        c[0] = (double)(-205.0f / 72.0f * nd);
        c[1] = (double)(8.0f / 5.0f);
        c[2] = (double)(-1.0f / 5.0f);
        c[3] = (double)(8.0f / 315.0f);
        c[4] = (double)(-1.0f / 560.0f);
        return 5;
    default:
        return 0;
    }
}

// s_integrate.cpp:308-334 with tau already advanced (:317)
StepScalars make_scalars(const axs_params &p, double tau)
{
    StepScalars s{};
    double dx = (double)p.space_step;
    s.dt = (double)p.time_step;
    double T0 = (double)p.T_initial;
    if (p.eom == AXS_EOM_HAMILTONIAN) {
        // s_integrate.cpp:364-380, tau already advanced
        double t2 = tau * tau;
        s.dcoef = s.dt / t2;
        s.hdt = (double)(0.5f * p.time_step);
        s.alpha = t2 / (dx * dx);
        s.cl = (-t2) * (double)p.lambda;
        s.tau2 = 1.0;                                  // the "- 1.0f" term
        s.T03 = (T0 * T0) / ((double)3.0f * t2);
    } else {
        s.alpha = (double)1.0f / (dx * dx);
        s.cl = ((double)(-1.0f) / (tau * tau)) * (double)p.lambda;
        s.tau2 = tau * tau;
        s.T03 = (T0 * T0) / (double)3.0f;
    }
    s.ncoef = stencil_coefficients(p.stencil_setting, p.ndims, s.c);
    return s;
}

StepScalarsF to_float(const StepScalars &s)
{
    StepScalarsF f{};
    f.dt = (float)s.dt;
    f.half_dt = 0.5f * f.dt;
    f.alpha = (float)s.alpha;
    f.cl = (float)s.cl;
    f.tau2 = (float)s.tau2;
    f.T03 = (float)s.T03;
    f.pot0 = (float)(s.T03 - s.tau2);
    f.ac0 = (float)(s.alpha * s.c[0]);
    f.dta = (float)(s.dt * s.alpha);
    f.dtcl = (float)(s.dt * s.cl);
    f.dtac0 = (float)(s.dt * s.alpha * s.c[0]);
    for (int i = 0; i < 5; ++i) f.c[i] = (float)s.c[i];
    f.ncoef = s.ncoef;
    return f;
}

int ensure_device(Context *c, void **buf, size_t *have, size_t bytes)
{
    if (bytes <= *have) return AXS_OK;
    if (*buf) {
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        AXS_HIP_TRY(hipFree(*buf));
        *buf = nullptr;
        *have = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    AXS_HIP_TRY(hipMalloc(buf, want));
    *have = want;
    return AXS_OK;
}

int ensure_scratch(Context *c, size_t bytes)
{
    return ensure_device(c, &c->scratch, &c->scratch_bytes, bytes);
}

int ensure_pinned(Context *c, size_t bytes)
{
    if (bytes <= c->pinned_bytes) return AXS_OK;
    if (c->pinned) {
        AXS_HIP_TRY(hipHostFree(c->pinned));
        c->pinned = nullptr;
        c->pinned_bytes = 0;
    }
    AXS_HIP_TRY(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
    c->pinned_bytes = bytes;
    return AXS_OK;
}

struct HostCopy {
    char *dst;
    const char *src;
};

static void host_copy_fn(size_t lo, size_t hi, void *arg)
{
    HostCopy *h = (HostCopy *)arg;
    memcpy(h->dst + lo, h->src + lo, hi - lo);
}

int copy_to_host(Context *c, void *dst, const void *src, size_t bytes)
{
    constexpr size_t kChunk = (size_t)64 << 20;
    if (bytes <= kChunk) {
        AXS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        return AXS_OK;
    }
    AXS_TRY(ensure_pinned(c, 2 * kChunk));
    for (int h = 0; h < 2; ++h)
        if (!c->stage_ev[h]) AXS_HIP_TRY(hipEventCreateWithFlags(&c->stage_ev[h],
                                                                 hipEventDisableTiming));
    char *stage[2] = {(char *)c->pinned, (char *)c->pinned + kChunk};
    const size_t n = (bytes + kChunk - 1) / kChunk;
    auto chunk_bytes = [&](size_t q) { return std::min(kChunk, bytes - q * kChunk); };
    // chunk q goes to stage q % 2; it is copied out after chunk q+1 is queued
    AXS_HIP_TRY(hipMemcpyAsync(stage[0], src, chunk_bytes(0), hipMemcpyDeviceToHost, c->stream));
    AXS_HIP_TRY(hipEventRecord(c->stage_ev[0], c->stream));
    for (size_t q = 0; q < n; ++q) {
        const int h = (int)(q & 1);
        if (q + 1 < n) {
            const size_t nb = chunk_bytes(q + 1);
            AXS_HIP_TRY(hipMemcpyAsync(stage[h ^ 1], (const char *)src + (q + 1) * kChunk, nb,
                                       hipMemcpyDeviceToHost, c->stream));
            AXS_HIP_TRY(hipEventRecord(c->stage_ev[h ^ 1], c->stream));
        }
        AXS_HIP_TRY(hipEventSynchronize(c->stage_ev[h]));
        HostCopy hc{(char *)dst + q * kChunk, stage[h]};
        parallel_for(chunk_bytes(q), host_copy_fn, &hc);
        // stage h is free again only after its copy-out: the transfer of
        // chunk q+2 into it is queued in the next iteration, after this one
    }
    return AXS_OK;
}

void parallel_for(size_t n, void (*fn)(size_t, size_t, void *), void *arg)
{
    unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = hw ? std::min(hw, 16u) : 4u;
    if (n < ((size_t)1 << 20) || nt <= 1) {
        fn(0, n, arg);
        return;
    }
    std::vector<std::thread> ts;
    size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        ts.emplace_back(fn, b, e, arg);
    }
    for (auto &t : ts) t.join();
}

// host-side (de)interleaving between the reference's separate double
// arrays and the device layouts
struct Pack64 {
    double *dst;              // pairs
    const double *a, *b;      // either may be null (zero)
};

static void pack64_fn(size_t lo, size_t hi, void *arg)
{
    Pack64 *c = (Pack64 *)arg;
    for (size_t i = lo; i < hi; ++i) {
        c->dst[2 * i] = c->a ? c->a[i] : 0.0;
        c->dst[2 * i + 1] = c->b ? c->b[i] : 0.0;
    }
}

struct Unpack64 {
    const double *src;        // pairs
    double *a, *b;            // either may be null (skip)
};

static void unpack64_fn(size_t lo, size_t hi, void *arg)
{
    Unpack64 *c = (Unpack64 *)arg;
    for (size_t i = lo; i < hi; ++i) {
        if (c->a) c->a[i] = c->src[2 * i];
        if (c->b) c->b[i] = c->src[2 * i + 1];
    }
}

// FP32 state quadruple (phi1, phi2, u1, u2), u = phidot + dt/2 ker (in double)
struct Pack32 {
    float *dst;
    const double *p1, *p2, *v1, *v2, *k1, *k2;
    double half_dt;
};

static void pack32_fn(size_t lo, size_t hi, void *arg)
{
    Pack32 *c = (Pack32 *)arg;
    for (size_t i = lo; i < hi; ++i) {
        double u1 = c->v1[i], u2 = c->v2[i];
        if (c->k1) u1 = u1 + c->half_dt * c->k1[i];
        if (c->k2) u2 = u2 + c->half_dt * c->k2[i];
        c->dst[4 * i] = (float)c->p1[i];
        c->dst[4 * i + 1] = (float)c->p2[i];
        c->dst[4 * i + 2] = (float)u1;
        c->dst[4 * i + 3] = (float)u2;
    }
}

struct PackPair32 {
    float *dst;
    const double *a, *b;
};

static void packpair32_fn(size_t lo, size_t hi, void *arg)
{
    PackPair32 *c = (PackPair32 *)arg;
    for (size_t i = lo; i < hi; ++i) {
        c->dst[2 * i] = c->a ? (float)c->a[i] : 0.0f;
        c->dst[2 * i + 1] = c->b ? (float)c->b[i] : 0.0f;
    }
}

struct Unpack32 {
    const float *src;
    int stride;               // 4 for the state, 2 for pairs
    double *a, *b;
};

static void unpack32_fn(size_t lo, size_t hi, void *arg)
{
    Unpack32 *c = (Unpack32 *)arg;
    for (size_t i = lo; i < hi; ++i) {
        if (c->a) c->a[i] = (double)c->src[c->stride * i];
        if (c->b) c->b[i] = (double)c->src[c->stride * i + 1];
    }
}

void destroy_plans(Context *c);
void destroy_inverse_plans(Context *c);

Context::~Context()
{
    if (stream) (void)hipStreamSynchronize(stream);
    destroy_snapshot(this);
    destroy_plans(this);
    destroy_inverse_plans(this);
    destroy_spec_plans(this);
    for (Member &m : members) {
        if (m.state_block) {
            (void)hipFree(m.state[0][0]);
        } else {
            for (int h = 0; h < 2; ++h)
                for (int a = 0; a < 3; ++a)
                    if (m.state[h][a]) (void)hipFree(m.state[h][a]);
        }
        if (m.aux_ker) (void)hipFree(m.aux_ker);
        if (m.aux_vel) (void)hipFree(m.aux_vel);
    }
    if (scratch) (void)hipFree(scratch);
    if (listbuf) (void)hipFree(listbuf);
    if (strbuf) (void)hipFree(strbuf);
    if (strtab) (void)hipFree(strtab);
    for (int h = 0; h < 2; ++h)
        if (batch_tab[h]) (void)hipFree(batch_tab[h]);
    for (MultiTables &t : mtab)
        for (int h = 0; h < 2; ++h)
            if (t.tab[h]) (void)hipFree(t.tab[h]);
    if (pinned) (void)hipHostFree(pinned);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (int h = 0; h < 2; ++h)
        if (stage_ev[h]) (void)hipEventDestroy(stage_ev[h]);
    if (stream) (void)hipStreamDestroy(stream);
}

int current_fields(Context *c, Member &m, FieldView *v)
{
    if (!m.has_state) {
        set_error("member has no state: call axs_set_member_state first");
        return AXS_ERR_STATE;
    }
    v->phi = m.state[c->cur][PHI];
    if (c->prec == AXS_FP64) {
        v->vel = m.state[c->cur][VEL];
        return AXS_OK;
    }
    AXS_TRY(fp32_refresh_aux(c, m));
    v->vel = m.aux_vel;
    return AXS_OK;
}

int fp32_refresh_aux(Context *c, Member &m)
{
    const bool ker_ok = ker_current(c, m);
    if (ker_ok && m.vel_valid) return AXS_OK;
    StepScalars s = make_scalars(c->p, c->tau);
    // a failed launch leaves both buffers unknown
    m.vel_valid = false;
    if (!ker_ok) m.ker_valid = false;
    AXS_TRY(launch_fp32_canonical(c, m, s, !ker_ok));
    if (!ker_ok) {
        m.ker_valid = true;
        m.ker_uploaded = false;
        m.ker_tau = c->tau;
    }
    m.vel_valid = true;
    return AXS_OK;
}

void state_changed(Context *c)
{
    for (Member &m : c->members) {
        m.ker_valid = false;
        m.ker_uploaded = false;
        m.vel_valid = false;
        m.version++;
    }
}

static int stencil_radius(int stencil) { return (stencil >= 0 && stencil <= 3) ? stencil + 1 : 0; }

static int validate_params(const axs_params *p, int N)
{
    if (p->ndims != 2 && p->ndims != 3) {
        set_error("ndims must be 2 or 3 (got %d)", p->ndims);
        return AXS_ERR_INVALID_ARG;
    }
    if (N < 4 || (N % 2) != 0) {     // s_run.cpp:32 asserts N % 2 == 0
        set_error("N must be even and >= 4 (got %d)", N);
        return AXS_ERR_INVALID_ARG;
    }
    int r = stencil_radius(p->stencil_setting);
    if (r == 0) {
        set_error("stencil_setting must be 0..3 (got %d)", p->stencil_setting);
        return AXS_ERR_UNSUPPORTED;
    }
    if (N < r) {
        set_error("N=%d smaller than the stencil radius %d", N, r);
        return AXS_ERR_INVALID_ARG;
    }
    double L = (p->ndims == 3) ? (double)N * N * N : (double)N * N;
    if (L >= 2147483648.0) {
        set_error("lattice of %.0f sites exceeds the 2^31 site limit", L);
        return AXS_ERR_UNSUPPORTED;
    }
    if (p->precision != AXS_FP64 && p->precision != AXS_FP32) {
        set_error("precision must be AXS_FP64 or AXS_FP32 (got %d)", p->precision);
        return AXS_ERR_INVALID_ARG;
    }
    if (!(p->time_step > 0.0f) || !(p->space_step > 0.0f)) {
        set_error("time_step and space_step must be positive");
        return AXS_ERR_INVALID_ARG;
    }
    if ((p->eom != AXS_EOM_FIELD_RESCALED && p->eom != AXS_EOM_HAMILTONIAN) ||
        (p->laplacian != AXS_LAPLACIAN_STENCIL && p->laplacian != AXS_LAPLACIAN_SPECTRAL)) {
        set_error("eom must be AXS_EOM_* and laplacian AXS_LAPLACIAN_* (got %d, %d)", p->eom,
                  p->laplacian);
        return AXS_ERR_INVALID_ARG;
    }
    if ((p->eom != AXS_EOM_FIELD_RESCALED || p->laplacian != AXS_LAPLACIAN_STENCIL) &&
        p->precision != AXS_FP64) {
        set_error("the Hamiltonian form and the spectral Laplacian need AXS_FP64");
        return AXS_ERR_UNSUPPORTED;
    }
    if (p->eom == AXS_EOM_HAMILTONIAN && p->laplacian == AXS_LAPLACIAN_SPECTRAL) {
        set_error("the reference has no Hamiltonian-form step with the spectral Laplacian");
        return AXS_ERR_UNSUPPORTED;
    }
    return AXS_OK;
}

// steps per pass of the higher-order stencils' multi-step kernels unless
// configured otherwise (measured best at 512^3 and 8192^2): 3D FP32 two
// (radius 4: one), FP64 one; 2D FP32 four, three, two at radius 2, 3, 4,
// FP64 two
static int high_order_default_depth(int nd, int prec, int stencil)
{
    if (nd == 2) return (prec == AXS_FP32) ? 5 - stencil : 2;
    return (prec == AXS_FP32 && stencil < 3) ? 2 : 1;
}

static int64_t total_sites(const Context *c)
{
    int64_t n = 0;
    for (const Member &m : c->members) n += m.L;
    return n;
}

static int create_impl(const axs_params *p, int n_members, const int *Ns, axs_ctx **out)
{
    if (!p || !out || n_members < 1 || !Ns) {
        set_error("axs_create: null argument or empty batch");
        return AXS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    for (int r = 0; r < n_members; ++r) AXS_TRY(validate_params(p, Ns[r]));

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device available");
        return AXS_ERR_NO_DEVICE;
    }
    if (p->device < 0 || p->device >= ndev) {
        set_error("device %d out of range (have %d)", p->device, ndev);
        return AXS_ERR_INVALID_ARG;
    }
    AXS_HIP_TRY(hipSetDevice(p->device));
    hipDeviceProp_t prop;
    AXS_HIP_TRY(hipGetDeviceProperties(&prop, p->device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", p->device,
                  prop.gcnArchName);
        return AXS_ERR_NO_DEVICE;
    }
    // the fixed workspace of axs_member_energy_noalloc on this device (once
    // per device); a failure here is reported by that entry point, which
    // tries again, and leaves the other entry points as they are
    (void)noalloc_setup(p->device);

    axs_ctx *h = new axs_ctx();
    Context *c = &h->impl;
    c->p = *p;
    c->nd = p->ndims;
    c->prec = p->precision;
    c->elem = (c->prec == AXS_FP64) ? 8 : 4;
    c->device = p->device;
    c->num_cu = prop.multiProcessorCount;
    c->tau = 1.0;                                              // physics.cpp:32
    if (const char *f = getenv("AXS_STEP_FAIL_AFTER")) c->step_fail_after = atoi(f);
    if (const char *f = getenv("AXS_PK_ROCFFT")) c->pk_rocfft = atoi(f) != 0;
    int rc = AXS_OK;
    auto fail = [&](int code) {
        delete h;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_error("stream/event creation failed");
        return fail(AXS_ERR_HIP);
    }
    c->members.resize(n_members);
    for (int r = 0; r < n_members; ++r) {
        Member &m = c->members[r];
        m.N = Ns[r];
        m.L = (c->nd == 3) ? (int64_t)m.N * m.N * m.N : (int64_t)m.N * m.N;
        // FP64: (phi, phidot, ker) pairs of doubles; FP32: one float
        // quadruple per site plus float-pair read-back buffers
        size_t bytes = (size_t)m.L * 16;
        const int narr = (c->prec == AXS_FP64) ? 3 : 1;
        if (c->prec == AXS_FP32) {
            // both halves in one allocation (Member::state_block), the
            // second on a 256-byte boundary
            const size_t stride = (bytes + 255) & ~(size_t)255;
            char *blk = nullptr;
            if (hipMalloc(reinterpret_cast<void **>(&blk), 2 * stride) != hipSuccess) {
                set_error("device allocation of %zu bytes failed (member %d)", 2 * stride, r);
                return fail(AXS_ERR_OUT_OF_MEMORY);
            }
            m.state_block = true;
            m.state[0][0] = blk;
            m.state[1][0] = blk + stride;
            if (hipMemsetAsync(blk, 0, 2 * stride, c->stream) != hipSuccess) {
                set_error("hipMemsetAsync failed");
                return fail(AXS_ERR_HIP);
            }
        } else {
            for (int hh = 0; hh < 2; ++hh)
                for (int a = 0; a < narr; ++a) {
                    if (hipMalloc(&m.state[hh][a], bytes) != hipSuccess) {
                        set_error("device allocation of %zu bytes failed (member %d)", bytes, r);
                        return fail(AXS_ERR_OUT_OF_MEMORY);
                    }
                    if (hipMemsetAsync(m.state[hh][a], 0, bytes, c->stream) != hipSuccess) {
                        set_error("hipMemsetAsync failed");
                        return fail(AXS_ERR_HIP);
                    }
                }
        }
        if (c->prec == AXS_FP32) {
            size_t pb = (size_t)m.L * 8;
            if (hipMalloc(&m.aux_ker, pb) != hipSuccess || hipMalloc(&m.aux_vel, pb) != hipSuccess) {
                set_error("device allocation of %zu bytes failed (member %d)", pb, r);
                return fail(AXS_ERR_OUT_OF_MEMORY);
            }
            (void)hipMemsetAsync(m.aux_ker, 0, pb, c->stream);
            (void)hipMemsetAsync(m.aux_vel, 0, pb, c->stream);
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("stream synchronize failed after allocation");
        return fail(AXS_ERR_HIP);
    }
    // several steps per pass unless configured otherwise (the fastest
    // shapes measured).  Standard stencil: FP32 3D four steps when the
    // lattices give four-step tiles chunks of 32 planes or more (192^3 and
    // up, ensembles), else three, 16-wave work-groups; FP64 3D three steps,
    // 8-wave work-groups, two rows per wave; 2D six steps (FP32 from 2^25
    // sites: eight).  Higher orders: high_order_default_depth.
    if (c->p.stencil_setting > 0)
        c->tune.depth = high_order_default_depth(c->nd, c->prec, c->p.stencil_setting);
    else if (c->nd == 2)   // FP32 from 2^25 sites on: eight (13% at 8192^2)
        c->tune.depth = (c->prec == AXS_FP32 && total_sites(c) >= (1LL << 25)) ? 8 : 6;
    else   // four steps, 16 waves of 2 rows: tiles of 24 x 56 sites
        c->tune.depth = (c->prec == AXS_FP32) ? (multi_chunk_cap(c, 24, 56, 4) >= 32 ? 4 : 3) : 3;
    if (c->nd == 3) c->tune.flags |= (c->prec == AXS_FP32) ? kTuneWG1024 : kTuneWG512;
    if (multi_depth(c) == 0) {
        c->tune.depth = 1;
        c->tune.flags &= ~(kTuneWG512 | kTuneWG1024);
    }
    (void)rc;
    *out = h;
    return AXS_OK;
}

static int check_member(axs_ctx *ctx, int member)
{
    if (!ctx) {
        set_error("null context");
        return AXS_ERR_INVALID_ARG;
    }
    if (member < 0 || member >= (int)ctx->impl.members.size()) {
        set_error("member %d out of range (batch of %d)", member, (int)ctx->impl.members.size());
        return AXS_ERR_INVALID_ARG;
    }
    AXS_HIP_TRY(hipSetDevice(ctx->impl.device));
    return AXS_OK;
}

}  // namespace axs

using namespace axs;

extern "C" {

int axs_create(const axs_params *p, axs_ctx **out)
{
    if (!p) {
        set_error("axs_create: null params");
        return AXS_ERR_INVALID_ARG;
    }
    int N = p->N;
    return create_impl(p, 1, &N, out);
}

int axs_create_batch(const axs_params *p, int n_members, const int *N_members, axs_ctx **out)
{
    return create_impl(p, n_members, N_members, out);
}

int axs_destroy(axs_ctx *ctx)
{
    if (!ctx) return AXS_OK;
    (void)hipSetDevice(ctx->impl.device);
    delete ctx;
    return AXS_OK;
}

int axs_num_members(const axs_ctx *ctx, int *n)
{
    if (!ctx || !n) {
        set_error("null argument");
        return AXS_ERR_INVALID_ARG;
    }
    *n = (int)ctx->impl.members.size();
    return AXS_OK;
}

int axs_set_member_state(axs_ctx *ctx, int member, const double *phi1, const double *phi2,
                         const double *phidot1, const double *phidot2, const double *ker1,
                         const double *ker2)
{
    AXS_TRY(check_member(ctx, member));
    if (!phi1 || !phi2 || !phidot1 || !phidot2) {
        set_error("phi1, phi2, phidot1 and phidot2 are required");
        return AXS_ERR_INVALID_ARG;
    }
    Context *c = &ctx->impl;
    Member &m = c->members[member];
    size_t L = (size_t)m.L;
    int h = c->cur;
    AXS_TRY(ensure_pinned(c, L * 16));
    if (c->prec == AXS_FP64) {
        double *st = (double *)c->pinned;
        const double *srcs[3][2] = {{phi1, phi2}, {phidot1, phidot2}, {ker1, ker2}};
        for (int a = 0; a < 3; ++a) {
            Pack64 pk{st, srcs[a][0], srcs[a][1]};
            parallel_for(L, pack64_fn, &pk);
            AXS_HIP_TRY(hipMemcpy(m.state[h][a], st, L * 16, hipMemcpyHostToDevice));
        }
    } else {
        // FP32 form: u = phidot + dt/2 ker in double, rounded once; ker kept
        // for the state read-back before the first step
        float *st = (float *)c->pinned;
        Pack32 pk{st, phi1, phi2, phidot1, phidot2, ker1, ker2, 0.5 * (double)c->p.time_step};
        parallel_for(L, pack32_fn, &pk);
        AXS_HIP_TRY(hipMemcpy(m.state[h][PHI], st, L * 16, hipMemcpyHostToDevice));
        PackPair32 kp{st, ker1, ker2};
        parallel_for(L, packpair32_fn, &kp);
        AXS_HIP_TRY(hipMemcpy(m.aux_ker, st, L * 8, hipMemcpyHostToDevice));
        m.ker_valid = true;
        m.ker_uploaded = true;
    }
    m.has_state = true;
    m.vel_valid = false;
    m.version++;
    return AXS_OK;
}

int axs_get_member_state(axs_ctx *ctx, int member, double *phi1, double *phi2, double *phidot1,
                         double *phidot2, double *ker1, double *ker2)
{
    AXS_TRY(check_member(ctx, member));
    Context *c = &ctx->impl;
    Member &m = c->members[member];
    if (!m.has_state) {
        set_error("member %d has no state", member);
        return AXS_ERR_STATE;
    }
    size_t L = (size_t)m.L;
    int h = c->cur;
    AXS_TRY(ensure_pinned(c, L * 16));
    if (c->prec == AXS_FP64) {
        double *st = (double *)c->pinned;
        double *dst[3][2] = {{phi1, phi2}, {phidot1, phidot2}, {ker1, ker2}};
        for (int a = 0; a < 3; ++a) {
            if (!dst[a][0] && !dst[a][1]) continue;
            AXS_HIP_TRY(hipMemcpyAsync(st, m.state[h][a], L * 16, hipMemcpyDeviceToHost,
                                       c->stream));
            AXS_HIP_TRY(hipStreamSynchronize(c->stream));
            Unpack64 up{st, dst[a][0], dst[a][1]};
            parallel_for(L, unpack64_fn, &up);
        }
        return AXS_OK;
    }
    FieldView fv;
    AXS_TRY(current_fields(c, m, &fv));      // rebuilds ker and phidot
    float *st = (float *)c->pinned;
    struct {
        const void *src;
        size_t bytes;
        int stride;
        double *a, *b;
    } jobs[3] = {{m.state[h][PHI], L * 16, 4, phi1, phi2},
                 {m.aux_vel, L * 8, 2, phidot1, phidot2},
                 {m.aux_ker, L * 8, 2, ker1, ker2}};
    for (auto &j : jobs) {
        if (!j.a && !j.b) continue;
        AXS_HIP_TRY(hipMemcpyAsync(st, j.src, j.bytes, hipMemcpyDeviceToHost, c->stream));
        AXS_HIP_TRY(hipStreamSynchronize(c->stream));
        Unpack32 up{st, j.stride, j.a, j.b};
        parallel_for(L, unpack32_fn, &up);
    }
    return AXS_OK;
}

int axs_init_member(axs_ctx *ctx, int member, int mode, uint64_t seed)
{
    AXS_TRY(check_member(ctx, member));
    if (mode != AXS_INIT_KINV && mode != AXS_INIT_THERMAL) {
        set_error("unknown init mode %d", mode);
        return AXS_ERR_INVALID_ARG;
    }
    return run_init(&ctx->impl, ctx->impl.members[member], mode, seed);
}

int axs_set_tau(axs_ctx *ctx, double tau)
{
    if (!ctx) {
        set_error("null context");
        return AXS_ERR_INVALID_ARG;
    }
    // FP32: a ker recomputed from phi (and the phidot rebuilt from it)
    // belongs to the tau it was computed at (ker_current); an uploaded one
    // does not depend on tau
    ctx->impl.tau = tau;
    return AXS_OK;
}

int axs_get_tau(const axs_ctx *ctx, double *tau)
{
    if (!ctx || !tau) {
        set_error("null argument");
        return AXS_ERR_INVALID_ARG;
    }
    *tau = ctx->impl.tau;
    return AXS_OK;
}

int axs_set_state(axs_ctx *ctx, const double *phi1, const double *phi2, const double *phidot1,
                  const double *phidot2, const double *ker1, const double *ker2, double tau)
{
    AXS_TRY(axs_set_member_state(ctx, 0, phi1, phi2, phidot1, phidot2, ker1, ker2));
    return axs_set_tau(ctx, tau);
}

int axs_get_state(axs_ctx *ctx, double *phi1, double *phi2, double *phidot1, double *phidot2,
                  double *ker1, double *ker2, double *tau)
{
    AXS_TRY(axs_get_member_state(ctx, 0, phi1, phi2, phidot1, phidot2, ker1, ker2));
    if (tau) *tau = ctx->impl.tau;
    return AXS_OK;
}

// the work-item tables of every pass depth that `nsteps` steps take
// (next_pass_depth), built before any graph capture
static int prepare_passes(Context *c, int64_t nsteps)
{
    const int depth = multi_depth(c);
    if (depth < 1) return AXS_OK;
    bool ready[kMaxDepth + 1] = {};
    for (int64_t left = nsteps; left > 0;) {
        const int d = next_pass_depth(c, left);
        if (d >= 1 && !ready[d]) {
            AXS_TRY(prepare_multi_batch(c, d));
            ready[d] = true;
        }
        // once a depth-D pass comes, depth-D passes follow down to the
        // remainder (next_pass_depth's deeper passes need more steps left)
        if (d == depth) {
            left %= depth;
            continue;
        }
        left -= (d >= 1) ? d : 1;
    }
    return AXS_OK;
}

int axs_step(axs_ctx *ctx, int64_t nsteps)
{
    AXS_TRY(check_member(ctx, 0));
    Context *c = &ctx->impl;
    if (nsteps < 0) {
        set_error("nsteps must be >= 0");
        return AXS_ERR_INVALID_ARG;
    }
    for (Member &m : c->members)
        if (!m.has_state) {
            set_error("every member needs a state before axs_step");
            return AXS_ERR_STATE;
        }
    if (nsteps == 0) return AXS_OK;
    AXS_TRY(prepare_batch(c));
    const double dt = (double)c->p.time_step;
    // the spectral step's rocFFT calls stay out of graph capture
    const bool graph = (c->tune.flags & kTuneGraph) && nsteps > 1 &&
                       c->p.laplacian == AXS_LAPLACIAN_STENCIL;
    AXS_TRY(prepare_passes(c, nsteps));
    const double tau0 = c->tau;
    const int cur0 = c->cur;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    // the events bracket the launches themselves, or the graph launch
    if (graph)
        AXS_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    else
        AXS_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    int rc = AXS_OK;
    int launches = 0;      // launches that returned AXS_OK (direct loop)
    for (int64_t n = 0; n < nsteps && rc == AXS_OK;) {
        double tau_next = c->tau + dt;                    // s_integrate.cpp:317
        StepScalars s = make_scalars(c->p, tau_next);
        const int64_t left = nsteps - n;
        const int d = next_pass_depth(c, left);
        if (!graph && c->step_fail_after >= 0 && launches >= c->step_fail_after) {
            set_error("axs_step: failure point after %d launches (AXS_STEP_FAIL_AFTER)",
                      launches);
            rc = AXS_ERR_HIP;
            break;
        }
        if (d >= 1) {
            StepScalars ss[kMaxDepth];
            double t = c->tau;
            for (int l = 0; l < d; ++l) {
                t = t + dt;                               // the same sequence of sums
                ss[l] = make_scalars(c->p, t);
            }
            rc = launch_multi_steps(c, ss, d);
            if (rc == AXS_OK) {
                c->tau = t;
                ++launches;
            }
            n += d;
        } else {
            rc = launch_step(c, s);
            if (rc == AXS_OK) {
                c->tau = tau_next;
                ++launches;
            }
            n += 1;
        }
    }
    if (graph) {
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc == AXS_OK && e == hipSuccess)
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (rc == AXS_OK && e != hipSuccess) {
            set_error("step graph capture failed: %s", hipGetErrorString(e));
            rc = AXS_ERR_HIP;
        }
    }
    auto drop_graph = [&]() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        ge = nullptr;
        g = nullptr;
    };
    if (rc != AXS_OK) {
        // a captured loop never ran: back to the state before the call; a
        // direct loop keeps the steps that were launched, and what was
        // derived from the state before them is dropped
        drop_graph();
        if (graph) {
            c->tau = tau0;
            c->cur = cur0;
        } else if (launches > 0) {
            state_changed(c);
        }
        return rc;
    }
    if (graph) {
        hipError_t e = hipEventRecord(c->ev0, c->stream);
        if (e == hipSuccess && c->step_fail_after >= 0) {
            set_error("axs_step: failure point at the graph launch (AXS_STEP_FAIL_AFTER)");
            e = hipErrorLaunchFailure;
        } else if (e == hipSuccess) {
            e = hipGraphLaunch(ge, c->stream);
            if (e != hipSuccess) set_error("hipGraphLaunch failed: %s", hipGetErrorString(e));
        } else {
            set_error("hipEventRecord failed: %s", hipGetErrorString(e));
        }
        if (e != hipSuccess) {
            // the graph did not run: the capture's tau and ping-pong index
            // are taken back
            drop_graph();
            c->tau = tau0;
            c->cur = cur0;
            return AXS_ERR_HIP;
        }
    }
    // the launches are queued: the state is the new one from here on, also
    // if waiting for them fails
    state_changed(c);
    hipError_t e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    drop_graph();
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (e != hipSuccess) {
        set_error("axs_step: waiting for the step launches failed: %s", hipGetErrorString(e));
        return AXS_ERR_HIP;
    }
    c->last_step_ms = (double)ms / (double)nsteps;
    return AXS_OK;
}

int axs_member_strings(axs_ctx *ctx, int member, int32_t *coords, int64_t cap, int64_t *count)
{
    AXS_TRY(check_member(ctx, member));
    if (!count || cap < 0) {
        set_error("count pointer required, cap >= 0");
        return AXS_ERR_INVALID_ARG;
    }
    return run_strings(&ctx->impl, ctx->impl.members[member], coords, cap, count, nullptr);
}

int axs_strings(axs_ctx *ctx, int32_t *coords, int64_t cap, int64_t *count)
{
    return axs_member_strings(ctx, 0, coords, cap, count);
}

int axs_member_axion(axs_ctx *ctx, int member, double *axion)
{
    AXS_TRY(check_member(ctx, member));
    if (!axion) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_axion(&ctx->impl, ctx->impl.members[member], axion);
}

int axs_string_counts(axs_ctx *ctx, int64_t *n_plaquettes, int64_t *num_cores)
{
    AXS_TRY(check_member(ctx, 0));
    return run_string_counts(&ctx->impl, n_plaquettes, num_cores);
}

int axs_member_background(axs_ctx *ctx, int member, double sums[6])
{
    AXS_TRY(check_member(ctx, member));
    if (!sums) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_background(&ctx->impl, ctx->impl.members[member], sums);
}

int axs_background(axs_ctx *ctx, double sums[6]) { return axs_member_background(ctx, 0, sums); }

int axs_member_nan_check(axs_ctx *ctx, int member, int64_t *first_nan)
{
    AXS_TRY(check_member(ctx, member));
    if (!first_nan) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_nan_check(&ctx->impl, ctx->impl.members[member], first_nan);
}

int axs_member_energy(axs_ctx *ctx, int member, double out[3])
{
    AXS_TRY(check_member(ctx, member));
    if (!out) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    if (ctx->impl.p.eom != AXS_EOM_FIELD_RESCALED) {
        set_error("the energy functional is defined for the field-rescaled form");
        return AXS_ERR_UNSUPPORTED;
    }
    return run_energy(&ctx->impl, ctx->impl.members[member], out);
}

int axs_energy(axs_ctx *ctx, double out[3]) { return axs_member_energy(ctx, 0, out); }

int axs_member_energy_noalloc(axs_ctx *ctx, int member, double out[3])
{
    AXS_TRY(check_member(ctx, member));
    if (!out) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    if (ctx->impl.p.eom != AXS_EOM_FIELD_RESCALED) {
        set_error("the energy functional is defined for the field-rescaled form");
        return AXS_ERR_UNSUPPORTED;
    }
    Workspace ws;
    AXS_TRY(noalloc_setup(ctx->impl.device));     // kept from creation when it succeeded there
    AXS_TRY(noalloc_workspace(ctx->impl.device, &ws));
    return run_energy(&ctx->impl, ctx->impl.members[member], out, &ws);
}

int axs_member_powerspec(axs_ctx *ctx, int member, int binning, int nbins, double *ks, double *pk,
                         int32_t *count)
{
    AXS_TRY(check_member(ctx, member));
    if (!ks || !pk || !count) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_powerspec(&ctx->impl, ctx->impl.members[member], binning, nbins, ks, pk, count);
}

int axs_member_powerspec_field(axs_ctx *ctx, int member, int field, int binning, int nbins,
                               double *ks, double *pk, int32_t *count)
{
    AXS_TRY(check_member(ctx, member));
    if (!ks || !pk || !count) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_powerspec_field(&ctx->impl, ctx->impl.members[member], field, binning, nbins, ks,
                               pk, count);
}

int axs_powerspec(axs_ctx *ctx, int binning, int nbins, double *ks, double *pk, int32_t *count)
{
    return axs_member_powerspec(ctx, 0, binning, nbins, ks, pk, count);
}

int axs_powerspec_all(axs_ctx *ctx, int binning, int nbins, double *ks, double *pk, int32_t *count)
{
    AXS_TRY(check_member(ctx, 0));
    if (!ks || !pk || !count) {
        set_error("null output");
        return AXS_ERR_INVALID_ARG;
    }
    return run_powerspec_all(&ctx->impl, binning, nbins, ks, pk, count);
}

int axs_set_kernel_config(axs_ctx *ctx, const axs_kernel_config *cfg)
{
    if (!ctx || !cfg) {
        set_error("null argument");
        return AXS_ERR_INVALID_ARG;
    }
    Context *c = &ctx->impl;
    const int depth = cfg->depth ? cfg->depth : 1;
    const int RAD = stencil_radius_of(c);
    const int wg = cfg->flags & (AXS_MARCH_WG512 | AXS_MARCH_WG1024);
    // the multi-step kernels' work-group size (3D): configured, else the
    // built one of the higher orders, else 4 waves
    const int W = (RAD > 1 && !wg) ? ((c->prec == AXS_FP64) ? 8 : 16) : multi_waves(cfg->flags);
    bool rows_ok;
    if (depth >= 2 && c->nd == 2) {
        // several steps per pass, 2D: one strip per wave
        rows_ok = cfg->rows == 0 && wg == 0;
        bool fits = cfg->kernel != AXS_KERNEL_SIMPLE && march2d_depth_built(depth, RAD, c->prec);
        for (const Member &m : c->members) fits = fits && march2d_supported(c, m.N);
        if (!fits) {
            set_error("depth %d in 2D needs the stencil Laplacian, sides >= 64, a non-simple "
                      "kernel and a built depth (stencil 0: 2, 3, 4, 6, 8; higher orders: 1-4 "
                      "FP32, 1-2 FP64)", depth);
            return AXS_ERR_UNSUPPORTED;
        }
    } else if (depth >= 2) {
        // several steps per pass, 3D, on every member
        const int R = cfg->rows ? cfg->rows
                                : (RAD > 1 ? high_order_rows(c->prec, RAD)
                                           : (W == 4 ? 4
                                                     : ((W == 8 && c->prec == AXS_FP32) ? 3 : 2)));
        rows_ok = wg != (AXS_MARCH_WG512 | AXS_MARCH_WG1024) &&
                  multi_shape_built(depth, W, R, c->prec, RAD);
        bool fits = cfg->kernel != AXS_KERNEL_SIMPLE && depth <= (RAD > 1 ? 2 : 5);
        for (const Member &m : c->members) fits = fits && march2_supported(c, m.N);
        if (!fits) {
            set_error("depth %d needs 3D, the stencil Laplacian, sides >= 64, a non-simple "
                      "kernel and depth <= 5 (stencil 0) or <= 2 (higher orders)", depth);
            return AXS_ERR_UNSUPPORTED;
        }
    } else if (RAD > 1) {
        // one step per pass: the multi-step kernel where it applies, with
        // its built shape (3D) or one strip per wave (2D)
        if (c->nd == 2)
            rows_ok = cfg->rows == 0 && wg == 0;
        else
            rows_ok = wg != (AXS_MARCH_WG512 | AXS_MARCH_WG1024) &&
                      multi_shape_built(1, W, cfg->rows ? cfg->rows : high_order_rows(c->prec, RAD),
                                        c->prec, RAD);
    } else {
        rows_ok = cfg->rows == 0 || cfg->rows == 4 || (cfg->rows == 8 && c->prec == AXS_FP32) ||
                  (cfg->rows == 2 && c->prec == AXS_FP64);
    }
    if (cfg->kernel < AXS_KERNEL_AUTO || cfg->kernel > AXS_KERNEL_MARCH || !rows_ok ||
        cfg->chunk < 0 || depth < 1 || depth > (c->nd == 2 ? 8 : 5) || (cfg->flags & ~63) != 0) {
        set_error("unsupported kernel configuration (kernel %d rows %d chunk %d depth %d "
                  "flags %d)", cfg->kernel, cfg->rows, cfg->chunk, cfg->depth, cfg->flags);
        return AXS_ERR_UNSUPPORTED;
    }
    for (const Member &m : c->members)
        if (cfg->chunk > 0 && (m.N % cfg->chunk) != 0) {
            set_error("chunk %d does not divide N=%d", cfg->chunk, m.N);
            return AXS_ERR_INVALID_ARG;
        }
    c->tune.kernel = cfg->kernel;
    if (depth >= 2 || RAD > 1)
        c->tune.rtb = cfg->rows;
    else if (c->prec == AXS_FP32)
        c->tune.rj32 = cfg->rows ? cfg->rows : 8;
    else
        c->tune.rj64 = cfg->rows ? cfg->rows : 4;
    c->tune.ci32 = cfg->chunk;
    c->tune.ci64 = cfg->chunk;
    c->tune.depth = depth;
    c->tune.flags = cfg->flags;
    return AXS_OK;
}

int axs_get_kernel_config(const axs_ctx *ctx, axs_kernel_config *cfg)
{
    if (!ctx || !cfg) {
        set_error("null argument");
        return AXS_ERR_INVALID_ARG;
    }
    const Context *c = &ctx->impl;
    cfg->kernel = c->tune.kernel;
    cfg->rows = (c->tune.depth >= 2 || c->p.stencil_setting > 0) ? c->tune.rtb
                                     : ((c->prec == AXS_FP32) ? c->tune.rj32 : c->tune.rj64);
    cfg->chunk = (c->prec == AXS_FP32) ? c->tune.ci32 : c->tune.ci64;
    cfg->depth = c->tune.depth;
    cfg->flags = c->tune.flags;
    return AXS_OK;
}

int axs_autotune(axs_ctx *ctx, int trials, axs_kernel_config *best, double *ms_per_step)
{
    AXS_TRY(check_member(ctx, 0));
    double ms = 0.0;
    AXS_TRY(run_autotune(&ctx->impl, trials, &ms));
    if (best) AXS_TRY(axs_get_kernel_config(ctx, best));
    if (ms_per_step) *ms_per_step = ms;
    return AXS_OK;
}

int axs_last_step_ms(const axs_ctx *ctx, double *ms)
{
    if (!ctx || !ms) {
        set_error("null argument");
        return AXS_ERR_INVALID_ARG;
    }
    *ms = ctx->impl.last_step_ms;
    return AXS_OK;
}

int axs_member_laplacian(axs_ctx *ctx, int member, int mode, double *lap1, double *lap2)
{
    AXS_TRY(check_member(ctx, member));
    if (mode != AXS_LAPLACIAN_STENCIL && mode != AXS_LAPLACIAN_SPECTRAL) {
        set_error("mode must be AXS_LAPLACIAN_STENCIL or AXS_LAPLACIAN_SPECTRAL");
        return AXS_ERR_INVALID_ARG;
    }
    return run_laplacian(&ctx->impl, ctx->impl.members[member], mode, lap1, lap2);
}

int axs_snapshot_begin(axs_ctx *ctx, int member, int fields, int64_t count)
{
    AXS_TRY(check_member(ctx, member));
    Member &m = ctx->impl.members[member];
    if (fields <= 0 || (fields & ~(AXS_SNAP_PHI | AXS_SNAP_VEL | AXS_SNAP_KER)) != 0) {
        set_error("fields must be a non-empty set of AXS_SNAP_* bits (got %d)", fields);
        return AXS_ERR_INVALID_ARG;
    }
    if (count < 0 || count > m.L) {
        set_error("count %lld outside [0, %lld]", (long long)count, (long long)m.L);
        return AXS_ERR_INVALID_ARG;
    }
    return snapshot_begin(&ctx->impl, m, fields, count ? count : m.L);
}

int axs_snapshot_end(axs_ctx *ctx, double *phi1, double *phi2, double *phidot1,
                     double *phidot2, double *ker1, double *ker2, double *tau)
{
    if (!ctx) {
        set_error("null context");
        return AXS_ERR_INVALID_ARG;
    }
    AXS_HIP_TRY(hipSetDevice(ctx->impl.device));      // may be another host thread
    double *out[6] = {phi1, phi2, phidot1, phidot2, ker1, ker2};
    return snapshot_end(&ctx->impl, out, tau);
}

int axs_bytes_per_site(const axs_ctx *ctx, double *scheme_bytes, double *alg_bytes)
{
    if (!ctx) {
        set_error("null context");
        return AXS_ERR_INVALID_ARG;
    }
    const Context *c = &ctx->impl;
    double w = (double)c->elem;
    if (alg_bytes) *alg_bytes = 2.0 * 4.0 * w;     // read + write phi1, phi2, vel1, vel2
    if (scheme_bytes) {   // FP64 also moves the two kernel arrays
        double b = (c->prec == AXS_FP64) ? 2.0 * 6.0 * w : 2.0 * 4.0 * w;
        *scheme_bytes = b / (double)multi_step_depth(c);
    }
    return AXS_OK;
}

const char *axs_strerror(int status)
{
    switch (status) {
    case AXS_OK: return "ok";
    case AXS_ERR_INVALID_ARG: return "invalid argument";
    case AXS_ERR_UNSUPPORTED: return "unsupported parameter";
    case AXS_ERR_OUT_OF_MEMORY: return "out of device memory";
    case AXS_ERR_HIP: return "HIP runtime error";
    case AXS_ERR_CAPACITY: return "output buffer too small";
    case AXS_ERR_NO_DEVICE: return "no usable gfx950 device";
    case AXS_ERR_FFT: return "rocFFT error";
    case AXS_ERR_STATE: return "call out of order";
    default: return "unknown status";
    }
}

const char *axs_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

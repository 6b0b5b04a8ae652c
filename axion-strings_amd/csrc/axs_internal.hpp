// axs_internal.hpp -- shared declarations of the gfx950 implementation of
// the Axion-Strings uniform-lattice path (context, device buffers, kernel
// launchers).  Host code and kernels are written for CDNA4 (64-wide waves).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "axs_gpu.h"

namespace axs {

// ---------------------------------------------------------------------------
// errors

void set_error(const char *fmt, ...);

#define AXS_HIP_TRY(expr)                                                     \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            ::axs::set_error("%s failed: %s (%s:%d)", #expr,                  \
                             hipGetErrorString(e_), __FILE__, __LINE__);      \
            return (e_ == hipErrorOutOfMemory) ? AXS_ERR_OUT_OF_MEMORY        \
                                               : AXS_ERR_HIP;                 \
        }                                                                     \
    } while (0)

#define AXS_TRY(expr)                                                         \
    do {                                                                      \
        int s_ = (expr);                                                      \
        if (s_ != AXS_OK) return s_;                                          \
    } while (0)

// ---------------------------------------------------------------------------
// per-step uniform scalars.  Computed once per step on the host in double,
// in the reference's expression order (s_integrate.cpp:308-334).

struct StepScalars {
    // Hamiltonian form (s_integrate.cpp:358-395): alpha = tau^2 / dx^2,
    // cl = -tau^2 * lambda, tau2 = 1, T03 = T0^2 / (3 tau^2), and the
    // drift phi + dcoef * (pi + hdt * ker) with dcoef = dt / tau^2,
    // hdt = 0.5f * dt (a float product)
    double dcoef, hdt;
    double dt;      // (double) parameters.time_step
    double alpha;   // 1.0f / pow_2(dx)
    double cl;      // (-1.0f / pow_2(tau)) * lambda, tau after the update
    double tau2;    // pow_2(tau)
    double T03;     // pow_2(T_initial) / 3.0f
    double c[5];    // stencil coefficients, s_integrate.cpp:9-14
    int ncoef;      // 2 for stencil_setting 0
};

struct StepScalarsF {
    float dt, alpha, cl, tau2, T03, half_dt;
    float pot0;     // T03 - tau2, rounded once from the double difference
    float ac0;      // alpha * c[0]
    // the step's kick scaled by dt (rounded once from the double products):
    // dta = dt alpha, dtcl = dt cl, dtac0 = dt alpha c[0]
    float dta, dtcl, dtac0;
    float c[5];
    int ncoef;
};

int stencil_coefficients(int stencil, int ndims, double *c);
StepScalars make_scalars(const axs_params &p, double tau_after);
StepScalarsF to_float(const StepScalars &s);

// most time steps one multi-step launch takes (2D strips: 8, 3D: 5)
constexpr int kMaxDepth = 8;

// ---------------------------------------------------------------------------
// one lattice of the context

// slots of Member::state.  FP64: three (x, y) double pairs per site, the
// reference's all_data arrays interleaved by component.  FP32: slot PHI holds
// (phi1, phi2, u1, u2) per site, u = phidot + dt/2 ker the half-step velocity.
enum ArrayId { PHI = 0, VEL = 1, KER = 2 };

struct Member {
    int N = 0;
    int64_t L = 0;
    void *state[2][3] = {{nullptr}};   // ping-pong halves
    // FP32: both halves in one allocation at state[0][0] (the second at L 16
    // bytes rounded up to 256): with separate allocations the halves'
    // relative placement in memory changed the 1024^3 four-step pass
    // between 1.70 and 1.98 ms per step from lattice to lattice; in one
    // allocation it ran 1.72-1.75 on 13 lattices of 14
    bool state_block = false;
    // FP32 only: (ker1, ker2) of the current state while ker_valid, and the
    // (phidot1, phidot2) read-back buffer, float pairs
    void *aux_ker = nullptr;
    void *aux_vel = nullptr;
    bool has_state = false;
    // FP32: aux_ker holds ker of the current state.  An uploaded ker (or the
    // zero of a fresh upload) belongs to the state whatever tau is; a ker
    // recomputed from phi belongs to the tau it was computed at (ker_tau).
    bool ker_valid = true;
    bool ker_uploaded = true;
    double ker_tau = 0.0;
    bool vel_valid = false;   // FP32: aux_vel holds phidot of the current state
    uint64_t version = 0;     // bumped whenever the state changes
    // background sums and energy components of state `samp_version` at tau
    // `samp_tau` (one pass computes both)
    bool samp_valid = false;
    uint64_t samp_version = 0;
    double samp_tau = 0.0;
    double samp[9] = {0.0};
};

struct FftPlan;

// one member of an ensemble as seen by the one-launch-per-step kernels
struct BatchMember {
    int N, nTj, nTk, CI;     // 3D march tiling (2D: N only)
    uint32_t item_begin;     // first work item: 3D march tile, 2D 256-site block
    uint32_t L;
    void *in[3];             // state[src] slots
    void *out[3];            // state[dst] slots
};

// march-kernel variant switches (Tune::flags)
constexpr int kTunePrefetch = 1;   // keep the next plane's loads in flight
constexpr int kTuneNT = 2;         // non-temporal stores of the new state
constexpr int kTuneJFast = 4;      // tiles ordered j fastest instead of k

struct Tune {
    int kernel = AXS_KERNEL_AUTO;
    int rj32 = 8;      // rows per work-item, FP32 march
    int rj64 = 4;      // rows per work-item, FP64 march
    int ci32 = 0;      // planes per chunk, 0 = automatic
    int ci64 = 0;
    int depth = 1;     // time steps per pass of the multi-step kernels (1: the
                       // one-step kernels for the 7/5-point stencil)
    int rtb = 0;       // rows per work-item of the multi-step kernel, 0 = by side
    int flags = kTunePrefetch | kTuneNT;
};

constexpr int kTuneGraph = 8;      // capture the step loop in a HIP graph
constexpr int kTuneWG512 = 16;     // multi-step kernel: 8 waves per work-group
constexpr int kTuneWG1024 = 32;    // multi-step kernel: 16 waves per work-group

struct Context {
    axs_params p{};
    int nd = 2;
    int prec = AXS_FP64;
    int elem = 8;                 // bytes per field element
    std::vector<Member> members;
    int cur = 0;                  // ping-pong index holding the current state
    double tau = 1.0;             // physics.cpp:32
    Tune tune;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};   // copy_to_host chunk completions
    double last_step_ms = 0.0;
    int device = 0;
    int num_cu = 256;
    // failure point for tests of the error paths: axs_step reports an error
    // in place of its (n+1)-th launch (direct loop) or of the graph launch
    // (AXS_STEP_GRAPH); -1 = off.  Read from AXS_STEP_FAIL_AFTER at creation.
    int step_fail_after = -1;
    // power spectra through rocFFT at every side (AXS_PK_ROCFFT=1 at
    // creation), for comparing the two transforms
    bool pk_rocfft = false;

    // device scratch for reductions / compaction (grown on demand; growing
    // discards the contents)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // device buffer for string coordinate lists (separate from scratch so
    // the compaction's counts survive its growth)
    void *listbuf = nullptr;
    size_t listbuf_bytes = 0;
    // pinned host staging
    void *pinned = nullptr;
    size_t pinned_bytes = 0;

    // asynchronous snapshot (axs_snapshot.hip): device staging copy made on
    // `stream`, device-to-host copy on `copy_stream`
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_staged = nullptr, snap_done = nullptr;
    void *snap_dev = nullptr;
    size_t snap_dev_bytes = 0;
    void *snap_host = nullptr;
    size_t snap_host_bytes = 0;
    std::atomic<int> snap_pending{0};
    int snap_fields = 0;
    int64_t snap_count = 0;
    double snap_tau = 0.0;

    std::vector<FftPlan *> fft_plans;
    // data-independent sums of the power-spectrum bin tables in use (|k|
    // numerators and mode counts per bin), kept per (side, dimensions,
    // binning, bins): computed once, by the binning kernel's geometry pass
    struct PkGeom {
        int N = 0, nd = 0, binning = 0, nbins = 0;
        std::vector<double> ks;
        std::vector<long long> cnt;
    };
    std::vector<PkGeom> pk_geom;
    std::vector<void *> inv_plans;   // device-init inverse transforms (axs_init.hip)
    std::vector<void *> spec_plans;  // spectral-Laplacian transforms (axs_spectral.hip)

    // ensemble tables (device), one per ping-pong parity, rebuilt when the
    // launch shape changes
    BatchMember *batch_tab[2] = {nullptr, nullptr};
    uint32_t batch_items = 0;
    int batch_key = -1;          // rows * 100000 + chunk the tables were built for
    bool batch_march = false;    // 3D ensemble on the march kernel
    // work-item tables of the multi-step kernels for each number of steps
    // per pass in use (the configured one and the shallower passes that
    // finish a step count), both parities
    struct MultiTables {
        int key = -1;
        BatchMember *tab[2] = {nullptr, nullptr};
        uint32_t items = 0;
    };
    MultiTables mtab[kMaxDepth + 1];
    // string finder results of the last call (ballot words, block counts,
    // offsets), reused while the member's state is unchanged
    void *strbuf = nullptr;
    size_t strbuf_bytes = 0;
    int str_member = -1;
    uint64_t str_version = 0;
    uint64_t str_total = 0;
    bool str_scanned = false;   // block offsets of the cached words computed
    // ensemble-wide 2D count: member table and block -> member map (built once)
    void *strtab = nullptr;
    uint32_t strtab_blocks = 0;

    ~Context();
};

int ensure_scratch(Context *c, size_t bytes);
int ensure_device(Context *c, void **buf, size_t *have, size_t bytes);
int ensure_pinned(Context *c, size_t bytes);
// device-to-host copy of `bytes` into a caller's (pageable) buffer through
// the pinned staging buffer, in chunks whose transfers overlap the host-side
// copies of the previous chunk; synchronous on return
int copy_to_host(Context *c, void *dst, const void *src, size_t bytes);

// host-side parallel loop over [0, n) in contiguous chunks (std::thread)
void parallel_for(size_t n, void (*fn)(size_t begin, size_t end, void *arg), void *arg);

// ---------------------------------------------------------------------------
// launchers (return AXS_OK or an error code)

// one step of every member; reads state[cur], writes state[cur ^ 1]
int launch_step(Context *c, const StepScalars &s);

// 3D 2.5D streaming step (stencil_setting 0, N a multiple of 64)
bool march3d_supported(const Context *c, int N);
int launch_march3d(Context *c, Member &m, int src, int dst, const StepScalars &s);
bool march2_supported(const Context *c, int N);
// steps per launch of the multi-step kernels for this configuration, or 0
// when the one-step kernels run (radius 1 at depth 1, simple kernel, ...)
int multi_depth(const Context *c);
// the same for D steps per pass instead of the configured number
int multi_depth_for(const Context *c, int D);
// steps per pass of the multi-step pass that finishes a remainder of r < D
// steps (the deepest usable one, at most r), or 0: one-step launches
int remainder_depth(const Context *c, int r);
int next_pass_depth(const Context *c, int64_t left);
// forget the work-item tables (the launch shape changed)
void invalidate_multi_tables(Context *c);
// time steps per launch the configuration gives every member (>= 1)
int multi_step_depth(const Context *c);
// the multi-step kernel is built for depth D, W waves, R rows per wave,
// form `prec` and stencil radius RAD
bool multi_shape_built(int D, int W, int R, int prec, int RAD);
int multi_waves(int flags);
int multi_waves_of(const Context *c);
int multi_rows(const Context *c, int N, int D, int W);
// chunk length cap of the 3D multi-step kernels for tiles TJ x TK
int multi_chunk_cap(const Context *c, int TJ, int TK, int H);
// largest divisor of N not above cap
inline int divisor_below(int N, int cap)
{
    int d = cap < N ? cap : N;
    while (N % d != 0) --d;
    return d;
}
int high_order_rows(int prec, int RAD);
int stencil_radius_of(const Context *c);
// higher-order stencils (radius 2-4), 3D (axs_march2_hi.hip)
int launch_multi_high_order(Context *c, int src, int dst, const StepScalars *s, int depth);
// `depth` steps (scalars s[0..depth-1] of tau+dt, tau+2dt, ...) in one pass
// per member
int launch_multi_steps(Context *c, const StepScalars *s, int depth);
// work-item tables of the multi-step kernel for an ensemble
int prepare_multi_batch(Context *c, int D);
// 2D multi-step kernel (axs_march2d.hip)
bool march2d_supported(const Context *c, int N);
bool march2d_depth_built(int D, int RAD, int prec);
int prepare_march2d_items(Context *c, int D);
int launch_march2d(Context *c, int src, const StepScalars *s, int depth);
// higher-order stencils (radius 2-4), 2D (axs_march2d_hi.hip)
int launch_march2d_high_order(Context *c, int src, const StepScalars *s, int depth, bool pf);
// field-rescaled step with the spectral Laplacian, one member
int run_spectral_step(Context *c, Member &m, int src, int dst, const StepScalars &s);
// Laplacian of the current field (AXS_LAPLACIAN_*) into host arrays
int run_laplacian(Context *c, Member &m, int mode, double *h1, double *h2);
void destroy_spec_plans(Context *c);
void destroy_snapshot(Context *c);
int snapshot_begin(Context *c, Member &m, int fields, int64_t count);
int snapshot_end(Context *c, double *out[6], double *tau);
// march geometry chosen for a member (rows per work-item, planes per chunk)
int march_rows(const Context *c);
int march_chunk(const Context *c, int N);
// ensemble launches (one per step for all members)
int launch_march3d_batch(Context *c, int src, const StepScalars &s);
int prepare_batch(Context *c);

// FP32: compute ker = F(phi, tau) and phidot = u - dt/2 ker into aux buffers
int launch_fp32_canonical(Context *c, Member &m, const StepScalars &s, bool recompute_ker);
// FP32: aux_ker holds ker of the current state at the context's tau
inline bool ker_current(const Context *c, const Member &m)
{
    return m.ker_valid && (m.ker_uploaded || m.ker_tau == c->tau);
}
// FP32: aux_ker and aux_vel rebuilt for the current state and tau (once per
// state and tau; a kept uploaded ker is not recomputed)
int fp32_refresh_aux(Context *c, Member &m);
// the state of every member has changed: drop what was derived from it
void state_changed(Context *c);

int run_string_counts(Context *c, int64_t *n_plaquettes, int64_t *num_cores);
// atan2(phi1, phi2) of every site of a member, rounded correctly, into a host array
int run_axion(Context *c, Member &m, double *out);
int run_strings(Context *c, Member &m, int32_t *coords, int64_t cap, int64_t *count,
                int64_t *num_cores);
// device workspace handed to a launcher in place of the context's scratch
struct Workspace {
    void *p = nullptr;
    size_t bytes = 0;
};
// the fixed per-GPU workspace of axs_member_energy_noalloc: its device
// address, taken once per device with the device current (axs_reduce.hip)
int noalloc_setup(int device);
int noalloc_workspace(int device, Workspace *ws);

int run_background(Context *c, Member &m, double out[6]);
int run_nan_check(Context *c, Member &m, int64_t *first);
int run_init(Context *c, Member &m, int mode, uint64_t seed);
int run_autotune(Context *c, int trials, double *best_ms);
// ws: a caller-given workspace used instead of the context's scratch
// (nothing is allocated); null: the context's scratch
int run_energy(Context *c, Member &m, double out[3], const Workspace *ws = nullptr);
int run_powerspec_all(Context *c, int binning, int nbins, double *ks, double *pk,
                      int32_t *count);
int run_powerspec(Context *c, Member &m, int binning, int nbins, double *ks, double *pk,
                  int32_t *count);
int run_powerspec_field(Context *c, Member &m, int field, int binning, int nbins, double *ks,
                        double *pk, int32_t *count);

// device pointers to the current fields of a member: phi (FP64 pairs or
// FP32 quadruples) and phidot (FP64 pairs or FP32 pairs) -- for the FP32 form
// this runs the kernel that rebuilds phidot (and ker) from u.
struct FieldView {
    const void *phi, *vel;
};
int current_fields(Context *c, Member &m, FieldView *v);

}  // namespace axs

struct axs_ctx {
    axs::Context impl;
};

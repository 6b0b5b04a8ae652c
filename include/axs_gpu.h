/*
 * axs_gpu.h -- C-ABI of the MI355X (gfx950) uniform-lattice evolution path.
 *
 * This is the drop-in boundary: the host driver (our counterpart of the
 * reference's src/standard/s_run.cpp) calls these functions where the
 * reference calls its in-process routines.  Every entry point returns an int
 * status (AXS_OK == 0); no C++ exception crosses it; every pointer argument
 * is a plain host pointer owned by the caller.  Host field arrays are double
 * in the reference layout (src/common/common.hpp:28, 81-119):
 *     2D  m = i + N*j            (i fastest)
 *     3D  m = (i*N + j)*N + k    (k fastest)
 *
 * A context owns its device memory and one HIP stream.  It is not thread-
 * safe: use one context per host thread.  Calls are synchronous with respect
 * to the host buffers they take (results are in place when the call returns).
 */
#ifndef AXS_GPU_H
#define AXS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define AXS_OK                 0
#define AXS_ERR_INVALID_ARG    1   /* bad pointer, size, index or parameter  */
#define AXS_ERR_UNSUPPORTED    2   /* parameter outside the supported set    */
#define AXS_ERR_OUT_OF_MEMORY  3   /* device allocation failed               */
#define AXS_ERR_HIP            4   /* a HIP runtime call failed              */
#define AXS_ERR_CAPACITY       5   /* output buffer too small (count is set) */
#define AXS_ERR_NO_DEVICE      6   /* no usable gfx950 device                */
#define AXS_ERR_FFT            7   /* rocFFT plan or execution failed        */
#define AXS_ERR_STATE          8   /* call out of order (e.g. no state yet)  */

/* ---- forms ------------------------------------------------------------- */
#define AXS_FP64  0   /* plain form: reference velocity-Verlet, bit-exact      */
#define AXS_FP32  1   /* phi and its velocity in FP32; tau and sums in FP64    */

#define AXS_INIT_KINV    0 /* complex GRF, |F| ~ 1/|k| below pi/(4dx), <|phi|^2>=1 */
#define AXS_INIT_THERMAL 1 /* thermal spectrum of gaussian_thermal (init.cpp)     */

#define AXS_BIN_LINEAR 0   /* reference bins floor(|k|)-1, N/2 of them          */
#define AXS_BIN_LOG    1   /* nbins log-spaced bins over 1 <= |k| < N/2          */

/*
 * Run parameters.  Mirrors the fields of the reference's `parameters` struct
 * that the evolution path reads (src/common/common.hpp:30-69), with the same
 * types (dt, dx, lambda and T_initial are float there and here), plus the
 * two build forms.  `N` is per lattice; a batch gives one N per member.
 */
typedef struct axs_params {
    int   ndims;            /* NDIMS: 2 or 3                        (common.hpp:38) */
    int   N;                /* points per side, even                (common.hpp:39) */
    float space_step;       /* dx                                   (common.hpp:40) */
    float time_step;        /* dt                                   (common.hpp:41) */
    float lambda;           /* quartic coupling                     (common.hpp:35) */
    float T_initial;        /* T0 in the thermal-mass term          (common.hpp:37) */
    int   stencil_setting;  /* 0..3, s_integrate.cpp:9-14           (common.hpp:42) */
    int   thr;              /* string threshold, percent            (common.hpp:62) */
    int   precision;        /* AXS_FP64 or AXS_FP32                                 */
    int   device;           /* HIP device ordinal                                   */
    int   eom;              /* AXS_EOM_* (0: the reference's default form)          */
    int   laplacian;        /* AXS_LAPLACIAN_* (0: the reference's default)         */
} axs_params;

/* Equation-of-motion form of the step.  The Hamiltonian form and the
 * spectral Laplacian are the reference's alternative code paths (selected
 * there by editing s_run.cpp:215-216 / s_integrate.cpp:321-329); both need
 * AXS_FP64, and the reference has no Hamiltonian step with the spectral
 * Laplacian.                                                             */
#define AXS_EOM_FIELD_RESCALED 0   /* vvsl_field_rescaled, s_integrate.cpp:305-348  */
#define AXS_EOM_HAMILTONIAN    1   /* vvsl_hamiltonian_form, s_integrate.cpp:358-395;
                                      the velocity arrays hold pi = phidot * a^2   */
#define AXS_LAPLACIAN_STENCIL  0   /* sparse stencil matrix, s_integrate.cpp:322-324 */
#define AXS_LAPLACIAN_SPECTRAL 1   /* compute_laplacian_fft, s_integrate.cpp:179-296,
                                      used at :326-328                             */

typedef struct axs_ctx axs_ctx;

/* ---- lifecycle --------------------------------------------------------- */

/* Replaces the allocation block of run_standard (s_run.cpp:35-58): the
 * stencil matrix of build_coefficient_matrix (s_integrate.cpp:5-149) is not
 * built -- the kernels apply the stencil directly.  One lattice.           */
int axs_create(const axs_params *p, axs_ctx **out);

/* An ensemble of n_members independent lattices of side N_members[r] that
 * share every other parameter and advance in lockstep (one tau).  No
 * reference counterpart: the reference ran ensembles as separate processes.*/
int axs_create_batch(const axs_params *p, int n_members, const int *N_members,
                     axs_ctx **out);

/* Frees everything (s_run.cpp:221-231). */
int axs_destroy(axs_ctx *ctx);

int axs_num_members(const axs_ctx *ctx, int *n_members);

/* ---- state ------------------------------------------------------------- */

/* Uploads the state of member `member` (0 for a single lattice): phi1, phi2,
 * phidot1, phidot2 and the current kernels ker1_curr, ker2_curr of the
 * reference's all_data (s_internal.hpp:6-21).  ker1/ker2 may be NULL, which
 * means zero -- the reference's calloc'd start (s_run.cpp:44-47).         */
int axs_set_member_state(axs_ctx *ctx, int member,
                         const double *phi1, const double *phi2,
                         const double *phidot1, const double *phidot2,
                         const double *ker1, const double *ker2);

/* Downloads the same six arrays (any pointer may be NULL to skip it). */
int axs_get_member_state(axs_ctx *ctx, int member,
                         double *phi1, double *phi2,
                         double *phidot1, double *phidot2,
                         double *ker1, double *ker2);

/* Generates initial conditions for member `member` directly in device
 * memory: AXS_INIT_KINV (the benchmark configurations' field, phidot = 0)
 * or AXS_INIT_THERMAL (the spectrum of gaussian_thermal, init.cpp:33-200,
 * replacing its MKL VSL normals and FFTW inverse transform by a Philox
 * counter-based generator and rocFFT).  ker is set to zero as after the
 * reference's calloc (s_run.cpp:44-47).  Same seed -> same field in the
 * FP64 and FP32 forms (rounded once).                                      */
int axs_init_member(axs_ctx *ctx, int member, int mode, uint64_t seed);

/* Program time tau (the global `tau`, physics.cpp:3, starts at 1.0). */
int axs_set_tau(axs_ctx *ctx, double tau);
int axs_get_tau(const axs_ctx *ctx, double *tau);

/* Single-lattice shorthands for member 0. */
int axs_set_state(axs_ctx *ctx, const double *phi1, const double *phi2,
                  const double *phidot1, const double *phidot2,
                  const double *ker1, const double *ker2, double tau);
int axs_get_state(axs_ctx *ctx, double *phi1, double *phi2,
                  double *phidot1, double *phidot2,
                  double *ker1, double *ker2, double *tau);

/* ---- the step ---------------------------------------------------------- */

/* nsteps calls of vvsl_field_rescaled (s_integrate.cpp:305-348) on every
 * member: drift, tau += dt, Laplacian, potential, kick.  In the FP64 form
 * the result is bit-identical to the reference expression trees.          */
int axs_step(axs_ctx *ctx, int64_t nsteps);

/* ---- strings ----------------------------------------------------------- */

/* The axion field of member `member`: axion[m] = atan2(phi1[m], phi2[m])
 * for every site (s_run.cpp:92-94, 187-189, the all_data axion array),
 * rounded correctly to double (libm's atan2 is within an ulp of this; which
 * last bit it returns depends on the library).  axion has N^ndims entries. */
int axs_member_axion(axs_ctx *ctx, int member, double *axion);

/* cores2 / cores3 (string_finding.cpp:3-86) on axion = atan2(phi1, phi2)
 * (s_run.cpp:92-94) of member `member`, with the correctly rounded angles
 * of axs_member_axion.  Writes up to `cap` (i,j) or
 * (i,j,k) int32 tuples in ascending site order (the reference's single-
 * thread order) to coords (may be NULL) and the number of flagged sites to
 * *count.  Returns AXS_ERR_CAPACITY (with *count set) if cap is too small. */
int axs_member_strings(axs_ctx *ctx, int member, int32_t *coords, int64_t cap,
                       int64_t *count);
int axs_strings(axs_ctx *ctx, int32_t *coords, int64_t cap, int64_t *count);

/* Flagged-site count and, in 2D, num_cores = n - (list-adjacent pairs)
 * (s_run.cpp:99-109); in 3D num_cores == n_plaquettes.  One value per
 * member, arrays of length n_members, from one synchronisation for the
 * whole ensemble.  For a single lattice (and the last member of a 3D
 * ensemble) the search stays cached while the state is unchanged, so an
 * axs_member_strings / axs_strings call for it next only writes the list. */
int axs_string_counts(axs_ctx *ctx, int64_t *n_plaquettes, int64_t *num_cores);

/* ---- reductions -------------------------------------------------------- */

/* The six background sums of s_run.cpp:139-147 (not yet divided by the
 * length): phi1, phi2, phidot1, phidot2, |atan2(phi1,phi2)|,
 * sqrt(phi1^2+phi2^2).  Compensated FP64.  sums has 6 entries.            */
int axs_member_background(axs_ctx *ctx, int member, double sums[6]);
int axs_background(axs_ctx *ctx, double sums[6]);

/* The reference's debug scan of the field (s_run.cpp:4-12; its call at
 * s_run.cpp:213, before each step, is commented out there): *first_nan =
 * the index (reference layout) of the first site whose phi1 or phi2 is NaN,
 * -1 if there is none.  The reference then prints "Error: NaN encountered
 * in solution vector." with the step and asserts; axs_main does the same
 * with the optional tag debug_nan 1.                                      */
int axs_member_nan_check(axs_ctx *ctx, int member, int64_t *first_nan);

/* Energy components (kinetic, gradient, potential) times the cell volume;
 * no reference counterpart.  Compensated FP64.
 * FP32 form: one pass computes the background sums and the energy of a
 * member's state together; the second of the two calls on the same state
 * and tau returns the kept result without touching the device.           */
int axs_member_energy(axs_ctx *ctx, int member, double out[3]);
int axs_energy(axs_ctx *ctx, double out[3]);

/* axs_member_energy without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * axs_member_energy, bit for bit; the partial sums go to a fixed device
 * workspace of 512 KiB per GPU (enough for every lattice: at most 2048
 * blocks of partials), held until the process ends, instead of the
 * context's scratch.  Calls on one GPU that overlap (several host threads
 * or contexts at once) share that workspace and race.  Returns
 * AXS_ERR_CAPACITY if a reduction would need more than the workspace.     */
int axs_member_energy_noalloc(axs_ctx *ctx, int member, double out[3]);

/* ---- power spectrum ---------------------------------------------------- */

/* output_powerspec (pk_estimation.cpp:37-214) of phi2 (s_run.cpp:206) of
 * member `member`: AXS_BIN_LINEAR gives the reference's N/2 unit-width bins
 * (nbins ignored; arrays need N/2 entries); AXS_BIN_LOG gives nbins log bins
 * with pk the mean |F|^2 per mode.  ks, pk, count: one entry per bin.  Empty
 * bins have count 0.  Transform by rocFFT, bins in compensated FP64.       */
int axs_member_powerspec(axs_ctx *ctx, int member, int binning, int nbins,
                         double *ks, double *pk, int32_t *count);
int axs_powerspec(axs_ctx *ctx, int binning, int nbins,
                  double *ks, double *pk, int32_t *count);

/* output_powerspec(re, im) of one field of member `member`
 * (pk_estimation.cpp:37-214 takes any real array and an optional imaginary
 * one; s_run.cpp:206 passes phi2 alone): AXS_FIELD_PHI2 is
 * axs_member_powerspec; AXS_FIELD_PHI1, AXS_FIELD_PHIDOT1, AXS_FIELD_PHIDOT2
 * a real component; AXS_FIELD_PHI and AXS_FIELD_PHIDOT the complex fields
 * phi1 + i phi2 and phidot1 + i phidot2 (pi in the Hamiltonian form).
 * Binning and outputs as axs_member_powerspec.                            */
#define AXS_FIELD_PHI2    0
#define AXS_FIELD_PHI1    1
#define AXS_FIELD_PHI     2
#define AXS_FIELD_PHIDOT2 3
#define AXS_FIELD_PHIDOT1 4
#define AXS_FIELD_PHIDOT  5
int axs_member_powerspec_field(axs_ctx *ctx, int member, int field, int binning, int nbins,
                               double *ks, double *pk, int32_t *count);

/* The spectra of every member in one call (ensembles have no reference
 * counterpart; each member's spectrum is the axs_member_powerspec one).
 * Members of one side share batched rocFFT transforms and one binning
 * launch, with one synchronisation per batch.  Member r's bins start at
 * entry sum_{q<r} nb_q of ks, pk and count, nb_q = N_q/2 (AXS_BIN_LINEAR)
 * or nbins (AXS_BIN_LOG).                                                  */
int axs_powerspec_all(axs_ctx *ctx, int binning, int nbins,
                      double *ks, double *pk, int32_t *count);

/* ---- kernel configuration ---------------------------------------------- */

#define AXS_KERNEL_AUTO   0   /* march kernel where it applies, else simple   */
#define AXS_KERNEL_SIMPLE 1   /* one work-item per site, neighbours via cache  */
#define AXS_KERNEL_MARCH  2   /* 2.5D streaming along the slow axis (3D)       */

/* Launch shape of the step kernels; no reference counterpart (the
 * reference's loops are OpenMP `parallel for`).  Zero fields mean "choose
 * automatically".  The same results come out of every configuration (the
 * FP64 form stays bit-exact).                                             */
#define AXS_MARCH_PREFETCH 1   /* next plane's loads in flight during compute */
#define AXS_MARCH_NT_STORE 2   /* non-temporal stores of the new state        */
#define AXS_MARCH_JFAST    4   /* tiles ordered j fastest (default k fastest) */
#define AXS_STEP_GRAPH     8   /* axs_step captures its launches in a HIP graph */
#define AXS_MARCH_WG512   16   /* 3D multi-step kernel: 8 waves per work-group
                                  (else 4; higher orders: the built size)  */
#define AXS_MARCH_WG1024  32   /* 3D multi-step kernel: 16 waves per work-group */

typedef struct axs_kernel_config {
    int kernel;   /* AXS_KERNEL_*                                           */
    int rows;     /* rows per work-item: FP32 4 or 8, FP64 2 or 4; depth >= 2:
                     4 or 6 (8 at depth 2; depth 4 needs 8 or 16 waves),
                     with AXS_MARCH_WG512 2, 3 or 4, with AXS_MARCH_WG1024
                     2 or 3.  Higher orders, 3D: FP32 2 (16 waves), FP64 2
                     or, at stencil_setting 1, 3 (8 waves)                 */
    int chunk;    /* planes marched per work-group (divides N)              */
    int depth;    /* time steps per pass over the lattice: 1 (0 = 1), or,
                     with stencil 0, the stencil Laplacian and every side
                     >= 64: 3D FP32 2 to 4, or 5 with 16 waves of 2 rows
                     (a new context starts at 4
                     when its lattices fill the chip in 32-plane chunks --
                     192^3 and up, ensembles -- else 3, with
                     AXS_MARCH_WG1024), 3D FP64 2 or 3 with AXS_MARCH_WG512
                     (starts at 3), 2D 2, 3, 4, 6 or 8 (starts at 6, FP32
                     from 2^25 sites at 8; rows 0, no work-group flags).  stencil_setting 1-3 (sides
                     >= 64, kernel not SIMPLE): the multi-step kernels
                     run from depth 1 -- 3D FP32 1 or 2 (stencil 3: 1), 3D
                     FP64 1, 2D FP32 1-4, 2D FP64 1 or 2; a new context
                     starts at 3D FP32 2 (stencil 3: 1), FP64 1, 2D FP32 4,
                     3, 2 (stencil 1, 2, 3), FP64 2.
                     Step counts that are not a multiple take a
                     remainder of one or two steps in as many passes of
                     depth + 1 where that depth is built (3D FP32, depth
                     4), else end with shallower multi-step passes where
                     built, else with one-step launches; every split gives
                     the same bits.                                         */
    int flags;    /* AXS_MARCH_* bits                                       */
} axs_kernel_config;

int axs_set_kernel_config(axs_ctx *ctx, const axs_kernel_config *cfg);
int axs_get_kernel_config(const axs_ctx *ctx, axs_kernel_config *cfg);

/* Autotuner: times every candidate launch shape of the step kernel on this
 * context's lattices (trial steps read the current state and write only the
 * spare half, so the state and tau are unchanged), keeps the fastest and
 * returns it in *best (may be NULL) with its time per step in *ms_per_step
 * (may be NULL).  Results are cached per (ndims, sides, form, device) for
 * the process and, if the environment variable AXS_TUNE_CACHE names a file,
 * across processes.  trials <= 0 picks a default.                        */
int axs_autotune(axs_ctx *ctx, int trials, axs_kernel_config *best, double *ms_per_step);

/* ---- diagnostics ------------------------------------------------------- */

/* Average device duration (ms) of the step kernels launched by the last
 * axs_step call, measured with HIP events on the context's stream.        */
int axs_last_step_ms(const axs_ctx *ctx, double *ms_per_step);

/* Bytes the configured step kernel must move per site-update (the scheme
 * bytes: FP64 96, FP32 32, divided by the steps per pass, as one pass reads
 * and writes the state once for `depth` updates) and the path's algorithmic
 * bytes per site-update (read+write of phi and its velocity: FP64 64, FP32
 * 32).                                                                    */
int axs_bytes_per_site(const axs_ctx *ctx, double *scheme_bytes, double *alg_bytes);

/* The Laplacian of one member's current field, per component, into host
 * arrays of length N^ndims (either may be NULL): mode AXS_LAPLACIAN_STENCIL
 * gives (1/dx^2) * (stencil matrix) * phi as s_integrate.cpp:322-324 builds
 * it, AXS_LAPLACIAN_SPECTRAL gives compute_laplacian_fft
 * (s_integrate.cpp:179-296) -- the reference's two operators side by side
 * for validation.                                                         */
int axs_member_laplacian(axs_ctx *ctx, int member, int mode, double *lap1, double *lap2);

/* ---- asynchronous snapshots --------------------------------------------- */

/* The reference's snapshots copy the fields to disk synchronously
 * (fileio.cpp:15-46, called at s_run.cpp:166-199).  axs_snapshot_begin
 * copies the requested fields of the first `count` sites (0 = all) of one
 * member's current state into device staging memory on the context's
 * stream and queues the device-to-host transfer on a second stream, then
 * returns: stepping can continue while the data travels.  axs_snapshot_end
 * waits for the transfer and fills the host arrays (length count; any may
 * be NULL) and *tau (the snapshot's tau); it may be called from another
 * host thread while axs_step runs on the context.  One snapshot may be in
 * flight per context (AXS_ERR_STATE otherwise).                          */
#define AXS_SNAP_PHI 1   /* phi1, phi2                                     */
#define AXS_SNAP_VEL 2   /* phidot1, phidot2 (pi in the Hamiltonian form)  */
#define AXS_SNAP_KER 4   /* ker1, ker2                                     */
int axs_snapshot_begin(axs_ctx *ctx, int member, int fields, int64_t count);
int axs_snapshot_end(axs_ctx *ctx, double *phi1, double *phi2, double *phidot1,
                     double *phidot2, double *ker1, double *ker2, double *tau);

const char *axs_strerror(int status);

/* Message of the last failure on this thread (empty if none). */
const char *axs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* AXS_GPU_H */

// run.cpp -- the host driver of the uniform-lattice path: the counterpart
// of the reference's run_standard (src/standard/s_run.cpp:26-232), its
// physics scalars (src/common/physics.cpp) and its output writers
// (src/common/fileio.cpp), calling the GPU path only through include/axs_gpu.h.
//
// Output files keep the reference's names and formats, so the reference's
// analysis notebook reads them unchanged:
//   time-series.csv       s_run.cpp:71-78, 87-152   (+ E columns if asked)
//   snapshot-timings.csv  s_run.cpp:63-69, 160-164
//   snapshot{n}-phi1/2    fileio.cpp:15-46 (raw float64; 3D: first N^2)
//   snapshot{n}-strings.csv fileio.cpp:83-113
//   snapshot{n}-pk.csv    fileio.cpp:64-81
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "axs_gpu.h"
#include "axs_host.h"
#include "params.hpp"

namespace {

// ---------------------------------------------------------------------------
// physics.cpp:14-74.  tau is held in double (the reference's data_t), the
// derived scalars are float as there.

struct Physics {
    float m_saxion_initial = 0.f;
    float T_initial = 0.f;
    float m_eff_squared = 0.f;
    float light_crossing_time = 0.f;
    double tau = 1.0;
};

void set_physics_variables(const parameters &p, Physics &ph)
{
    ph.m_saxion_initial = sqrtf(p.lambda);                                     // :18
    ph.T_initial = p.T_initial;                                                // :23
    ph.m_eff_squared =
        (float)((double)p.lambda * ((double)ph.T_initial * ph.T_initial / 3.0 - 1.0));  // :26
    ph.light_crossing_time = 0.5f * p.N * p.space_step;                        // :30
    ph.tau = 1.0f;                                                             // :33
}

float physical_time(const Physics &ph) { return (float)(0.5f * (ph.tau * ph.tau)); }   // :38-40
float scale_factor(const Physics &ph) { return (float)ph.tau; }                        // :42-45
float hubble_parameter(const Physics &ph) { return (float)(1.0f / (ph.tau * ph.tau)); } // :51-53
float string_tension(const Physics &ph)                                                  // :65-67
{
    return (float)log(ph.m_saxion_initial / hubble_parameter(ph));
}

int final_step_of(int N, float dx, float dt)
{
    // s_run.cpp:82 with light_crossing_time = 0.5f * N * dx (float)
    float lct = 0.5f * N * dx;
    return (int)(roundf(lct / dt) - roundf(dx / dt) + 1);
}

float xi_of(int ndims, int64_t count, double tau, int N, float dx)
{
    float t_phys = (float)(0.5f * (tau * tau));
    float a = (float)tau;
    if (ndims == 2) {                                                   // s_run.cpp:111-114
        double side = (double)(N * dx * a);
        float physical_area = (float)(side * side);
        return (float)((double)((int)count * ((double)t_phys * t_phys)) / physical_area);
    }
    float total_length = (int)count * 2.0f / 3.0f;                      // s_run.cpp:123
    float physical_length = total_length * a;
    double side = (double)(N * dx * a);
    float physical_volume = (float)(side * side * side);
    return (float)((double)physical_length * ((double)t_phys * t_phys) / physical_volume);
}

// ---------------------------------------------------------------------------
// fileio.cpp.  Paths are built with snprintf into a std::string (the
// reference's fixed 209-byte alloca overflows for long directories).

struct Out {
    FILE *main = stdout, *ts = nullptr, *snap = nullptr;
    std::string dir;
    bool own_main = false;
};

std::string path_in(const Out &o, const char *name) { return o.dir + "/" + name; }

bool write_raw_file(const std::string &path, const double *data, size_t count)
{
    FILE *fp = fopen(path.c_str(), "w");
    if (!fp) return false;
    size_t w = fwrite(data, sizeof(double), count, fp);
    return fclose(fp) == 0 && w == count;
}

// Field snapshots leave the device asynchronously (axs_snapshot_begin) and
// are written by this thread while the main thread keeps stepping; at most
// one is in flight.  Joined before the next snapshot and before the
// context goes away.
struct SnapshotWriter {
    std::thread t;
    std::atomic<bool> failed{false};
    bool join()
    {
        if (t.joinable()) t.join();
        return !failed.load();
    }
    void start(axs_ctx *ctx, size_t count, std::string path1, std::string path2)
    {
        t = std::thread([this, ctx, count, path1, path2]() {
            std::vector<double> f1(count), f2(count);
            if (axs_snapshot_end(ctx, f1.data(), f2.data(), nullptr, nullptr, nullptr, nullptr,
                                 nullptr) != AXS_OK ||
                !write_raw_file(path1, f1.data(), count) || !write_raw_file(path2, f2.data(), count))
                failed.store(true);
        });
    }
    ~SnapshotWriter() { join(); }
};

bool save_strings(Out &o, const char *name, const int32_t *c, int64_t n, int nd)
{
    std::string path = path_in(o, name);
    fprintf(o.main, "  Saving string positions... at %s\n", path.c_str());
    FILE *fp = fopen(path.c_str(), "w");
    if (!fp) return false;
    for (int64_t i = 0; i < n; ++i) {
        if (nd == 2)
            fprintf(fp, "%d, %d\n", c[2 * i], c[2 * i + 1]);
        else
            fprintf(fp, "%d, %d, %d\n", c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    }
    fclose(fp);
    return true;
}

bool save_pk(Out &o, const char *name, const double *pk, const double *ks, const int32_t *count,
             int n_bins)
{
    std::string path = path_in(o, name);
    fprintf(o.main, "  Saving data... at %s\n", path.c_str());
    FILE *fp = fopen(path.c_str(), "w");
    if (!fp) return false;
    fprintf(fp, "k,pk,n_modes\n");
    for (int i = 0; i < n_bins; i++)
        if (count[i] != 0) fprintf(fp, "%e,%e,%d\n", ks[i], pk[i], count[i]);
    fclose(fp);
    return true;
}

bool open_streams(const parameters &p, Out &o)
{
    o.dir = p.output_directory;
    if (p.write_output_file) {                                           // fileio.cpp:118-123
        o.main = fopen(p.output_file_path, "w");
        if (!o.main) return false;
        o.own_main = true;
    }
    if (p.save_snapshots) {
        std::string path = path_in(o, "snapshot-timings.csv");
        o.snap = fopen(path.c_str(), "w");
        if (!o.snap) return false;
    }
    if (p.sample_time_series) {
        o.ts = fopen(p.ts_output_path, "w");
        if (!o.ts) return false;
    }
    return true;
}

void close_streams(Out &o)
{
    if (o.own_main && o.main) fclose(o.main);
    if (o.snap) fclose(o.snap);
    if (o.ts) fclose(o.ts);
}

// s_run.cpp:14-24.  n == 1 divides by zero in the reference; here it
// samples only step 0.
bool on_cadence(int enabled, int tstep, int n, int final_step)
{
    if (!enabled) return false;
    if (tstep == 0) return true;
    if (n < 2) return false;
    int every = final_step / (n - 1);
    return every > 0 && tstep % every == 0;
}

// ---------------------------------------------------------------------------
// full-state checkpoint: phi, phidot, ker (both components) and tau

const char kCkptMagic[8] = {'A', 'X', 'S', 'C', 'K', 'P', 'T', '1'};

struct CkptHeader {
    char magic[8];
    int32_t ndims, N;
    int64_t step;
    double tau;
};

bool write_checkpoint(const char *path, int nd, int N, int64_t step, double tau,
                      const std::vector<double> *arrays)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return false;
    CkptHeader h;
    memcpy(h.magic, kCkptMagic, 8);
    h.ndims = nd;
    h.N = N;
    h.step = step;
    h.tau = tau;
    bool ok = fwrite(&h, sizeof h, 1, fp) == 1;
    for (int a = 0; a < 6 && ok; ++a)
        ok = fwrite(arrays[a].data(), sizeof(double), arrays[a].size(), fp) == arrays[a].size();
    fclose(fp);
    return ok;
}

bool read_checkpoint(const char *path, int nd, int N, int64_t *step, double *tau,
                     std::vector<double> *arrays)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return false;
    CkptHeader h;
    bool ok = fread(&h, sizeof h, 1, fp) == 1 && memcmp(h.magic, kCkptMagic, 8) == 0 &&
              h.ndims == nd && h.N == N;
    for (int a = 0; a < 6 && ok; ++a)
        ok = fread(arrays[a].data(), sizeof(double), arrays[a].size(), fp) == arrays[a].size();
    fclose(fp);
    if (ok) {
        *step = h.step;
        *tau = h.tau;
    }
    return ok;
}

#define CHECK(call)                                                                          \
    do {                                                                                     \
        int rc_ = (call);                                                                    \
        if (rc_ != AXS_OK) {                                                                 \
            fprintf(stderr, "axs: %s failed: %s (%s)\n", #call, axs_strerror(rc_),           \
                    axs_last_error());                                                       \
            return 1;                                                                        \
        }                                                                                    \
    } while (0)

int run_standard(const parameters &p, Out &o)
{
    // s_run.cpp:32-33
    if (p.N % 2 != 0 || (p.NDIMS != 2 && p.NDIMS != 3)) {
        fprintf(stderr, "axs: N must be even and NDIMS 2 or 3\n");
        return 1;
    }
    const int nd = p.NDIMS, N = p.N;
    const int64_t length = (nd == 3) ? (int64_t)N * N * N : (int64_t)N * N;

    Physics ph;
    set_physics_variables(p, ph);

    axs_params ap{};
    ap.ndims = nd;
    ap.N = N;
    ap.space_step = p.space_step;
    ap.time_step = p.time_step;
    ap.lambda = p.lambda;
    ap.T_initial = p.T_initial;
    ap.stencil_setting = p.stencil_setting;
    ap.thr = p.thr;
    ap.precision = (strcmp(p.precision, "fp32") == 0) ? AXS_FP32 : AXS_FP64;
    ap.device = p.device;
    if (strcmp(p.precision, "fp32") != 0 && strcmp(p.precision, "fp64") != 0) {
        fprintf(stderr, "axs: precision must be fp64 or fp32\n");
        return 1;
    }
    const int init_mode = (strcmp(p.init_mode, "kinv") == 0) ? AXS_INIT_KINV : AXS_INIT_THERMAL;
    if (strcmp(p.init_mode, "kinv") != 0 && strcmp(p.init_mode, "thermal") != 0) {
        fprintf(stderr, "axs: init_mode must be thermal or kinv\n");
        return 1;
    }
    if (strcmp(p.eom, "field_rescaled") != 0 && strcmp(p.eom, "hamiltonian") != 0) {
        fprintf(stderr, "axs: eom must be field_rescaled or hamiltonian\n");
        return 1;
    }
    ap.eom = (strcmp(p.eom, "hamiltonian") == 0) ? AXS_EOM_HAMILTONIAN : AXS_EOM_FIELD_RESCALED;
    if (strcmp(p.laplacian, "stencil") != 0 && strcmp(p.laplacian, "spectral") != 0) {
        fprintf(stderr, "axs: laplacian must be stencil or spectral\n");
        return 1;
    }
    ap.laplacian =
        (strcmp(p.laplacian, "spectral") == 0) ? AXS_LAPLACIAN_SPECTRAL : AXS_LAPLACIAN_STENCIL;
    const bool log_bins = strcmp(p.pk_binning, "log") == 0;

    axs_ctx *ctx = nullptr;
    CHECK(axs_create(&ap, &ctx));
    struct Guard {
        axs_ctx *c;
        ~Guard() { axs_destroy(c); }
    } guard{ctx};
    SnapshotWriter writer;      // destroyed (joined) before the context

    // initial state: checkpoint, device generator or the host generator
    int first_step = 0;
    {
        std::vector<double> a[6];
        if (p.restart_path[0]) {
            for (auto &v : a) v.assign((size_t)length, 0.0);
            int64_t st = 0;
            double tau = 1.0;
            if (!read_checkpoint(p.restart_path, nd, N, &st, &tau, a)) {
                fprintf(stderr, "axs: cannot read checkpoint %s\n", p.restart_path);
                return 1;
            }
            CHECK(axs_set_state(ctx, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                a[4].data(), a[5].data(), tau));
            ph.tau = tau;
            first_step = (int)st;
        } else if (p.init_on_device) {
            CHECK(axs_init_member(ctx, 0, init_mode, p.seed));
            CHECK(axs_set_tau(ctx, 1.0));
        } else {
            for (int i = 0; i < 4; ++i) a[i].assign((size_t)length, 0.0);
            int rc = axs_host_init_fields(init_mode == AXS_INIT_KINV ? 0 : 1, nd, N, p.space_step,
                                          p.lambda, p.T_initial, p.seed, a[0].data(), a[1].data(),
                                          a[2].data(), a[3].data());
            if (rc != 0) {
                fprintf(stderr, "axs: host initial conditions failed (%d)\n", rc);
                return 1;
            }
            // ker_curr starts at zero (s_run.cpp:44-47)
            CHECK(axs_set_state(ctx, a[0].data(), a[1].data(), a[2].data(), a[3].data(), nullptr,
                                nullptr, 1.0));
        }
    }

    if (p.save_snapshots) fprintf(o.snap, "snapshot,tau,hubble_scale,string_tension,\n");
    if (p.sample_time_series) {                                           // s_run.cpp:71-78
        fprintf(o.ts, "time,");
        fprintf(o.ts, "scalefactor,");
        if (p.sample_strings) fprintf(o.ts, "xi,");
        if (p.sample_strings) fprintf(o.ts, "n_plaquettes,");
        if (p.sample_background)
            fprintf(o.ts, "phi1_bar,phi2_bar,phidot1_bar,phidot2_bar,axion_bar,saxion_bar");
        if (p.sample_energy) fprintf(o.ts, "%sE_kin,E_grad,E_pot", p.sample_background ? "," : "");
        fprintf(o.ts, "\n");
    }

    const int final_step = (p.n_steps > 0) ? p.n_steps
                                           : final_step_of(N, p.space_step, p.time_step);
    int n_snapshots_written = 0;
    std::vector<int32_t> coords;
    double step_ms_total = 0.0, sample_s = 0.0;
    int64_t steps_done = 0;

    // the reference's debug scan before a step (s_run.cpp:4-12, at :213):
    // the message, then the run stops (the reference asserts)
    auto nan_scan = [&](int step) {
        if (!p.debug_nan) return true;
        int64_t first = -1;
        if (axs_member_nan_check(ctx, 0, &first) != AXS_OK) {
            fprintf(stderr, "axs: axs_member_nan_check failed: %s\n", axs_last_error());
            return false;
        }
        if (first < 0) return true;
        fprintf(o.main, "Error: NaN encountered in solution vector.\n");
        fprintf(o.main, " tstep = %d\n", step);
        fflush(o.main);
        return false;
    };

    int tstep = first_step;
    while (tstep < final_step) {
        // next step at which something is sampled (sampling happens before
        // that step, s_run.cpp:85, 155)
        int next = tstep;
        while (next < final_step && !on_cadence(p.sample_time_series, next, p.n_samples, final_step) &&
               !on_cadence(p.save_snapshots, next, p.n_snapshots, final_step))
            ++next;
        if (next > tstep && p.debug_nan) {
            for (; tstep < next; ++tstep) {          // one step at a time, scanned before each
                if (!nan_scan(tstep)) return 1;
                CHECK(axs_step(ctx, 1));
                double ms = 0.0;
                axs_last_step_ms(ctx, &ms);
                step_ms_total += ms;
                steps_done += 1;
            }
            CHECK(axs_get_tau(ctx, &ph.tau));
        } else if (next > tstep) {
            CHECK(axs_step(ctx, next - tstep));
            double ms = 0.0;
            axs_last_step_ms(ctx, &ms);
            step_ms_total += ms * (next - tstep);
            steps_done += next - tstep;
            tstep = next;
            CHECK(axs_get_tau(ctx, &ph.tau));
        }
        if (tstep >= final_step) break;
        auto t0 = std::chrono::steady_clock::now();

        if (on_cadence(p.sample_time_series, tstep, p.n_samples, final_step)) {
            fprintf(o.ts, "%f", physical_time(ph));
            fprintf(o.ts, ",%f", scale_factor(ph));
            if (p.sample_strings) {
                int64_t n_plaq = 0, num_cores = 0;
                CHECK(axs_string_counts(ctx, &n_plaq, &num_cores));
                float xi = xi_of(nd, nd == 2 ? num_cores : n_plaq, ph.tau, N, p.space_step);
                fprintf(o.ts, ",%f", xi);
                fprintf(o.ts, ",%d", (int)n_plaq);
            }
            if (p.sample_background) {
                double s[6];
                CHECK(axs_background(ctx, s));
                fprintf(o.ts, ",%f,%f,%f,%f,%f,%f", s[0] / length, s[1] / length, s[2] / length,
                        s[3] / length, s[4] / length, s[5] / length);
            }
            if (p.sample_energy) {
                double e[3];
                CHECK(axs_energy(ctx, e));
                fprintf(o.ts, ",%e,%e,%e", e[0], e[1], e[2]);
            }
            fprintf(o.ts, "\n");
        }

        if (on_cadence(p.save_snapshots, tstep, p.n_snapshots, final_step)) {
            fprintf(o.main, "Writing snapshot %d:\n", n_snapshots_written);
            fprintf(o.snap, "%d,", n_snapshots_written);                   // s_run.cpp:160-164
            fprintf(o.snap, "%f,", ph.tau);
            fprintf(o.snap, "%f,", 1.0f / hubble_parameter(ph));
            fprintf(o.snap, "%f,", (p.lambda != 0.0f) ? string_tension(ph) : 0.0f);
            fprintf(o.snap, "\n");
            char name[64];
            if (p.save_fields) {
                if (!writer.join()) {
                    fprintf(stderr, "axs: writing a field snapshot failed\n");
                    return 1;
                }
                // 3D: the reference writes only the first N^2 values (fileio.cpp:31-46)
                size_t cnt = (nd == 2 || p.save_full_fields) ? (size_t)length : (size_t)N * N;
                CHECK(axs_snapshot_begin(ctx, 0, AXS_SNAP_PHI, (int64_t)cnt));
                snprintf(name, sizeof name, "snapshot%d-phi1", n_snapshots_written);
                std::string path1 = path_in(o, name);
                snprintf(name, sizeof name, "snapshot%d-phi2", n_snapshots_written);
                std::string path2 = path_in(o, name);
                fprintf(o.main, "  Saving data... at %s\n", path1.c_str());
                fprintf(o.main, "  Saving data... at %s\n", path2.c_str());
                writer.start(ctx, cnt, path1, path2);
            }
            if (p.save_strings) {
                int64_t n = 0;
                CHECK(axs_strings(ctx, nullptr, 0, &n));
                coords.assign((size_t)(n > 0 ? n : 1) * nd, 0);
                CHECK(axs_strings(ctx, coords.data(), n, &n));
                snprintf(name, sizeof name, "snapshot%d-strings.csv", n_snapshots_written);
                if (!save_strings(o, name, coords.data(), n, nd)) return 1;
            }
            if (p.save_pk) {
                int nb = log_bins ? p.pk_nbins : N / 2;
                std::vector<double> ks(nb), pk(nb);
                std::vector<int32_t> cnt(nb);
                CHECK(axs_powerspec(ctx, log_bins ? AXS_BIN_LOG : AXS_BIN_LINEAR, nb, ks.data(),
                                    pk.data(), cnt.data()));
                snprintf(name, sizeof name, "snapshot%d-pk.csv", n_snapshots_written);
                if (!save_pk(o, name, pk.data(), ks.data(), cnt.data(), nb)) return 1;
            }
            n_snapshots_written++;
        }
        sample_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        if (!nan_scan(tstep)) return 1;                                   // s_run.cpp:213
        CHECK(axs_step(ctx, 1));                                          // s_run.cpp:215
        double ms = 0.0;
        axs_last_step_ms(ctx, &ms);
        step_ms_total += ms;
        steps_done += 1;
        ++tstep;
        CHECK(axs_get_tau(ctx, &ph.tau));
    }

    if (!writer.join()) {
        fprintf(stderr, "axs: writing a field snapshot failed\n");
        return 1;
    }
    if (p.checkpoint_path[0]) {
        std::vector<double> a[6];
        for (auto &v : a) v.assign((size_t)length, 0.0);
        double tau = 0.0;
        CHECK(axs_get_state(ctx, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                            a[5].data(), &tau));
        if (!write_checkpoint(p.checkpoint_path, nd, N, final_step, tau, a)) {
            fprintf(stderr, "axs: cannot write checkpoint %s\n", p.checkpoint_path);
            return 1;
        }
    }
    if (steps_done > 0)
        fprintf(o.main, "Steps: %lld, device time per step %.4f ms, sampling %.3f s\n",
                (long long)steps_done, step_ms_total / (double)steps_done, sample_s);
    return 0;
}

void json_str(std::string &s, const char *k, const char *v)
{
    s += "\"";
    s += k;
    s += "\": \"";
    s += v;
    s += "\", ";
}

void json_num(std::string &s, const char *k, double v)
{
    char b[64];
    snprintf(b, sizeof b, "\"%s\": %.17g, ", k, v);
    s += b;
}

}  // namespace

extern "C" int axs_host_final_step(int N, float space_step, float time_step)
{
    return final_step_of(N, space_step, time_step);
}

extern "C" float axs_host_xi(int ndims, int64_t count, double tau, int N, float space_step)
{
    return xi_of(ndims, count, tau, N, space_step);
}

extern "C" int axs_host_params_json(const char *path, char *buf, int cap)
{
    parameters p;
    if (read_parameter_file(path, &p) != 0) return 1;
    std::string s = "{";
    json_num(s, "lambda", p.lambda);
    json_num(s, "enable_PRS", p.enable_PRS);
    json_num(s, "T_initial", p.T_initial);
    json_num(s, "NDIMS", p.NDIMS);
    json_num(s, "N", p.N);
    json_num(s, "space_step", p.space_step);
    json_num(s, "time_step", p.time_step);
    json_num(s, "stencil_setting", p.stencil_setting);
    json_num(s, "seed", p.seed);
    json_num(s, "write_output_file", p.write_output_file);
    json_str(s, "output_file_path", p.output_file_path);
    json_num(s, "save_snapshots", p.save_snapshots);
    json_num(s, "n_snapshots", p.n_snapshots);
    json_str(s, "output_directory", p.output_directory);
    json_num(s, "save_fields", p.save_fields);
    json_num(s, "save_strings", p.save_strings);
    json_num(s, "save_pk", p.save_pk);
    json_num(s, "sample_time_series", p.sample_time_series);
    json_num(s, "n_samples", p.n_samples);
    json_str(s, "ts_output_path", p.ts_output_path);
    json_num(s, "sample_strings", p.sample_strings);
    json_num(s, "sample_background", p.sample_background);
    json_num(s, "thr", p.thr);
    json_num(s, "enable_amr", p.enable_amr);
    json_num(s, "refinement_threshold", p.refinement_threshold);
    json_num(s, "init_from_snapshot", p.init_from_snapshot);
    json_num(s, "tau_initial", p.tau_initial);
    json_num(s, "n_steps", p.n_steps);
    json_str(s, "precision", p.precision);
    json_str(s, "init_mode", p.init_mode);
    json_num(s, "init_on_device", p.init_on_device);
    json_num(s, "sample_energy", p.sample_energy);
    json_str(s, "pk_binning", p.pk_binning);
    json_num(s, "pk_nbins", p.pk_nbins);
    json_num(s, "save_full_fields", p.save_full_fields);
    json_num(s, "device", p.device);
    json_str(s, "checkpoint_path", p.checkpoint_path);
    json_str(s, "restart_path", p.restart_path);
    json_str(s, "eom", p.eom);
    json_str(s, "laplacian", p.laplacian);
    json_num(s, "debug_nan", p.debug_nan);
    s.resize(s.size() - 2);
    s += "}";
    if ((int)s.size() + 1 > cap) return 2;
    memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

extern "C" int axs_host_run(const char *path)
{
    parameters p;
    if (read_parameter_file(path, &p) != 0) return 1;
    if (p.enable_amr) {
        fprintf(stderr, "axs: the AMR driver (enable_amr 1) is not part of this build\n");
        return 1;
    }
    Out o;
    if (!open_streams(p, o)) {
        fprintf(stderr, "axs: cannot open an output file\n");
        close_streams(o);
        return 1;
    }
    fprintf(o.main, "Running with %s\n", path);                              // main.cpp:15
    auto t0 = std::chrono::steady_clock::now();
    int rc = run_standard(p, o);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(o.main, "Time taken: %f seconds\n", secs);                         // main.cpp:22
    close_streams(o);
    return rc;
}

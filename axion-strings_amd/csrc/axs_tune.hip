// axs_tune.hip -- autotuner of the step kernel's launch shape.
//
// Candidates: for 3D lattices the march kernel with rows per work-item,
// planes per chunk and the prefetch / non-temporal-store switches, plus the
// one-site-per-work-item kernel; for every lattice, with and without a HIP
// graph around the step loop.  A trial launches the step repeatedly on the
// current state, writing only the spare ping-pong half and restoring the
// state index, so tuning never advances the simulation.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "axs_internal.hpp"

namespace axs {

namespace {

std::mutex g_tune_mu;
std::map<std::string, std::pair<Tune, double>> g_tune_cache;

std::string tune_key(const Context *c)
{
    hipDeviceProp_t prop;
    std::string dev = "gfx";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) dev = prop.gcnArchName;
    std::string k = dev + " nd" + std::to_string(c->nd) + (c->prec == AXS_FP32 ? " fp32" : " fp64") +
                    " st" + std::to_string(c->p.stencil_setting) + " eom" +
                    std::to_string(c->p.eom) + " lap" + std::to_string(c->p.laplacian) + " N";
    std::vector<int> Ns;
    for (const Member &m : c->members) Ns.push_back(m.N);
    std::sort(Ns.begin(), Ns.end());
    for (size_t i = 0; i < Ns.size(); ++i) {
        // run-length: "128x16" for sixteen members of side 128
        size_t j = i;
        while (j + 1 < Ns.size() && Ns[j + 1] == Ns[i]) ++j;
        k += (i ? "," : "") + std::to_string(Ns[i]) + "x" + std::to_string(j - i + 1);
        i = j;
    }
    return k;
}

void load_file_cache()
{
    const char *path = getenv("AXS_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    char line[1024];
    while (fgets(line, sizeof line, f)) {
        // key | kernel rj32 rj64 ci32 ci64 flags | ms
        char *bar1 = strchr(line, '|');
        if (!bar1) continue;
        char *bar2 = strchr(bar1 + 1, '|');
        if (!bar2) continue;
        std::string key(line, bar1 - line);
        while (!key.empty() && key.back() == ' ') key.pop_back();
        Tune t;
        double ms = 0;
        if (sscanf(bar1 + 1, "%d %d %d %d %d %d %d %d", &t.kernel, &t.rj32, &t.rj64, &t.ci32,
                   &t.ci64, &t.flags, &t.depth, &t.rtb) == 8 &&
            sscanf(bar2 + 1, "%lf", &ms) == 1)
            g_tune_cache[key] = {t, ms};
    }
    fclose(f);
}

void append_file_cache(const std::string &key, const Tune &t, double ms)
{
    const char *path = getenv("AXS_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s | %d %d %d %d %d %d %d %d | %.6f\n", key.c_str(), t.kernel, t.rj32, t.rj64,
            t.ci32, t.ci64, t.flags, t.depth, t.rtb, ms);
    fclose(f);
}

// time `trials` steps of the current configuration without changing state
int time_trial(Context *c, int trials, double *ms)
{
    AXS_TRY(prepare_batch(c));
    const int cur0 = c->cur;
    const double dt = (double)c->p.time_step;
    const int depth = multi_depth(c);     // 0: the one-step kernels
    StepScalars ss[kMaxDepth];
    double t = c->tau;
    for (int l = 0; l < kMaxDepth; ++l) {
        t = t + dt;
        ss[l] = make_scalars(c->p, t);
    }
    const StepScalars &s = ss[0];
    // one trial = one launch of one or several steps
    auto one = [&]() { return depth >= 1 ? launch_multi_steps(c, ss, depth) : launch_step(c, s); };
    AXS_TRY(one());                             // warm-up (code objects, caches)
    c->cur = cur0;
    const bool graph = (c->tune.flags & kTuneGraph) != 0;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (graph) {
        AXS_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = AXS_OK;
        for (int t = 0; t < trials && rc == AXS_OK; ++t) {
            rc = one();
            c->cur = cur0;
        }
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != AXS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        AXS_HIP_TRY(e);
        AXS_HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    }
    AXS_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    if (graph) {
        AXS_HIP_TRY(hipGraphLaunch(ge, c->stream));
    } else {
        for (int t = 0; t < trials; ++t) {
            AXS_TRY(one());
            c->cur = cur0;
        }
    }
    AXS_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    AXS_HIP_TRY(hipEventSynchronize(c->ev1));
    float f = 0.f;
    AXS_HIP_TRY(hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    *ms = (double)f / trials / (depth >= 1 ? depth : 1);
    return AXS_OK;
}

}  // namespace

int run_autotune(Context *c, int trials, double *best_ms)
{
    const std::string key = tune_key(c);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        static bool loaded = false;
        if (!loaded) {
            load_file_cache();
            loaded = true;
        }
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            c->tune = it->second.first;
            if (best_ms) *best_ms = it->second.second;
            return AXS_OK;
        }
    }
    if (trials <= 0) trials = 8;
    for (Member &m : c->members)
        if (!m.has_state) {
            set_error("autotune needs a state on every member");
            return AXS_ERR_STATE;
        }

    std::vector<Tune> cands;
    int minN = 1 << 30;
    for (const Member &m : c->members) minN = std::min(minN, m.N);
    const int RAD = stencil_radius_of(c);
    const bool march_ok = c->nd == 3 && RAD == 1;
    bool hi_ok = c->nd == 3 && RAD > 1;
    for (const Member &m : c->members) hi_ok = hi_ok && march2_supported(c, m.N);
    bool strips_ok = c->nd == 2;
    for (const Member &m : c->members) strips_ok = strips_ok && march2d_supported(c, m.N);
    for (int graph = 0; graph < 2; ++graph) {
        Tune t;
        t.kernel = AXS_KERNEL_SIMPLE;
        t.flags = graph ? kTuneGraph : 0;
        cands.push_back(t);
        if (strips_ok) {
            // 2D: several steps per pass, one strip per wave
            const int depths[8] = {1, 2, 3, 4, 5, 6, 7, 8};
            const int cj_opts[4] = {0, 32, 64, 128};
            for (int d : depths)
                for (int pf = 0; pf < 2; ++pf)
                    for (int ci = 0; ci < 4; ++ci) {
                        bool ok = true;
                        for (const Member &m : c->members)
                            ok = ok && (cj_opts[ci] == 0 || m.N % cj_opts[ci] == 0);
                        if (!ok) continue;
                        if (!march2d_depth_built(d, RAD, c->prec)) continue;
                        Tune u;
                        u.kernel = AXS_KERNEL_MARCH;
                        u.depth = d;
                        u.ci32 = u.ci64 = cj_opts[ci];
                        u.flags = (pf ? kTunePrefetch : 0) | kTuneNT | (graph ? kTuneGraph : 0);
                        cands.push_back(u);
                    }
        }
        if (hi_ok) {
            // higher-order stencils, 3D: the multi-step kernel at one or two
            // steps per pass in its built shape
            const int W = (c->prec == AXS_FP64) ? 8 : 16;
            const int chunk_opts3[3] = {0, 32, 64};
            for (int d = 1; d <= 2; ++d)
                for (int pf = 0; pf < 2; ++pf)
                    for (int ci = 0; ci < 3; ++ci) {
                        int chunk = chunk_opts3[ci];
                        bool ok = true;
                        for (const Member &m : c->members)
                            ok = ok && (chunk == 0 || m.N % chunk == 0);
                        if (!ok) continue;
                        Tune u;
                        u.kernel = AXS_KERNEL_MARCH;
                        u.depth = d;
                        u.ci32 = u.ci64 = chunk;
                        u.flags = (pf ? kTunePrefetch : 0) | kTuneNT |
                                  (W == 8 ? kTuneWG512 : kTuneWG1024) | (graph ? kTuneGraph : 0);
                        cands.push_back(u);
                    }
        }
        if (!march_ok) continue;
        const int rows_opts32[2] = {4, 8}, rows_opts64[2] = {2, 4};
        const int flag_opts[3] = {kTunePrefetch | kTuneNT, kTuneNT, kTunePrefetch};
        const int chunk_opts[4] = {0, 32, 64, 128};
        for (int ri = 0; ri < 2; ++ri)
            for (int fi = 0; fi < 3; ++fi)
                for (int ci = 0; ci < 4; ++ci) {
                    int chunk = chunk_opts[ci];
                    bool ok = true;
                    for (const Member &m : c->members) ok = ok && (chunk == 0 || m.N % chunk == 0);
                    if (!ok) continue;
                    Tune u;
                    u.kernel = AXS_KERNEL_MARCH;
                    u.rj32 = rows_opts32[ri];
                    u.rj64 = rows_opts64[ri];
                    u.ci32 = u.ci64 = chunk;
                    u.flags = flag_opts[fi] | (graph ? kTuneGraph : 0);
                    cands.push_back(u);
                }
        // several steps per pass: (depth, waves per work-group, rows per wave)
        const int shapes32[][3] = {{2, 4, 4}, {2, 8, 2}, {2, 8, 3}, {3, 4, 4}, {3, 8, 2},
                                   {3, 8, 3}, {3, 8, 4}, {3, 16, 2}, {4, 8, 3}, {4, 8, 4},
                                   {4, 16, 2}, {4, 16, 3}, {5, 16, 2}};
        const int shapes64[][3] = {{2, 8, 2}, {2, 8, 3}, {3, 8, 2}, {3, 8, 3}};
        std::vector<std::array<int, 3>> shapes;
        if (c->prec == AXS_FP32)
            for (const auto &x : shapes32) shapes.push_back({x[0], x[1], x[2]});
        else
            for (const auto &x : shapes64) shapes.push_back({x[0], x[1], x[2]});
        const int chunk_opts2[3] = {0, 32, 64};
        for (const auto &sh : shapes)
            for (int pf = 0; pf < 2; ++pf)
                for (int ci = 0; ci < 3; ++ci) {
                    int chunk = chunk_opts2[ci];
                    bool ok = true;
                    for (const Member &m : c->members) ok = ok && (chunk == 0 || m.N % chunk == 0);
                    if (!ok) continue;
                    Tune u;
                    u.kernel = AXS_KERNEL_MARCH;
                    u.depth = sh[0];
                    u.rtb = sh[2];
                    u.ci32 = u.ci64 = chunk;
                    u.flags = (pf ? kTunePrefetch : 0) | kTuneNT | (sh[1] == 8 ? kTuneWG512 : 0) |
                              (sh[1] == 16 ? kTuneWG1024 : 0) | (graph ? kTuneGraph : 0);
                    cands.push_back(u);
                }
    }
    const Tune saved = c->tune;
    Tune best = saved;
    double best_t = 1e300;
    for (const Tune &t : cands) {
        c->tune = t;
        if (t.kernel == AXS_KERNEL_MARCH && (t.depth >= 2 || RAD > 1)) {
            if (multi_depth(c) != t.depth) continue;
        } else if (t.kernel == AXS_KERNEL_MARCH) {
            bool ok = true;
            for (const Member &m : c->members) ok = ok && march3d_supported(c, m.N);
            if (!ok) continue;
        }
        double ms = 0.0;
        int rc = time_trial(c, trials, &ms);
        if (rc != AXS_OK) {
            c->tune = saved;
            c->batch_key = -1;
            invalidate_multi_tables(c);
            return rc;
        }
        if (ms < best_t) {
            best_t = ms;
            best = t;
        }
    }
    c->tune = best;
    c->batch_key = -1;          // tables follow the chosen shape
    invalidate_multi_tables(c);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {best, best_t};
        append_file_cache(key, best, best_t);
    }
    if (best_ms) *best_ms = best_t;
    (void)minN;
    return AXS_OK;
}

}  // namespace axs

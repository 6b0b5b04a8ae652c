#!/usr/bin/env python3
"""Lattice site-updates/s of the velocity-Verlet step (the reference's
vvsl_field_rescaled, src/standard/s_integrate.cpp:305-348) on one MI355X.

Workload: 3D 1024^3 periodic lattice, the FP32 form (phi and its velocity
in FP32, tau and every reduction in FP64) -- the headline -- beside the FP64
form on the same k^-1 Gaussian random field (seed 1998, generated in device
memory), lambda = 1, dx = 0.5, dt = 0.1 (float), T_initial = 4, 7-point
Laplacian.  One "step" advances the whole lattice by one time step; W warm-up
steps are followed by exactly K timed steps, with the stream synchronised
before and after.

The JSON line also carries
  roofline     : the step kernel against HBM.  One launch of the default
                 FP32 kernel advances D time steps (D = 4 at 1024^3) and
                 reads and writes the state once: its compulsory bytes are
                 32 B per site per launch (read + write of phi1, phi2, u1,
                 u2 in float).  `achieved` = 32 B x N^3 / average launch
                 duration (HIP events on the library's stream), `frac` =
                 achieved / 8 TB/s.  `site_update_rate` is the same bytes
                 counted once per site-update (32 B x D x N^3 per launch),
                 which exceeds the peak by design; `traffic` is the HBM
                 bytes one launch moves, from the committed rocprofv3 PMC
                 profile of the same kernel configuration (`traffic_source`
                 names it; null when the run's configuration differs);
  cpu_baseline : the CPU oracle (reference-structured FP64 step: drift,
                 two CSR sparse-matrix Laplacians, potential, kick, copy, with
                 OpenMP) timed on this host over a bounded sample;
  config3      : the 3D 1024^3 configuration as a whole run in both forms
                 -- always 1000 steps from the same seeded field, the string
                 count and the 64-bin log power spectrum every 50 steps --
                 with its accuracy gate: FP32 vs FP64 field rel-L2 <= 1e-4
                 and string count within 1% after the 1000 steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "axion-strings_amd"))

HBM_PEAK_GBS = 8000.0
# measured HBM bytes per launch of the FP32 step kernels at 1024^3, by time
# steps per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated); each file
# names the kernel configuration it was measured on
TRAFFIC_FILES = {1: os.path.join(ROOT, "profiles", "traffic_march_fp32_3d_1024.json"),
                 4: os.path.join(ROOT, "profiles", "traffic_multi_step_fp32_3d_1024.json"),
                 5: os.path.join(ROOT, "profiles", "traffic_multi_step_fp32_3d_1024_d5.json")}
# compulsory bytes per site and launch: the state read once and written once
PASS_BYTES = {"fp32": 32.0, "fp64": 96.0}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def gpu_leg(np, axs, N, precision, warmup, steps, T0, depth=0):
    lat = axs.Lattice(3, N, precision=precision, T_initial=T0)
    # device initial conditions: the first call builds rocFFT's plan, the
    # second is the steady cost (the field it leaves is the same)
    t0 = time.perf_counter()
    lat.init("kinv", seed=1998)
    init_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat.init("kinv", seed=1998)
    init_again = time.perf_counter() - t0
    lat.tau = 1.0
    if depth:
        try:
            # multi-step depths keep the context's work-group size
            flags = lat.kernel_config()["flags"] if depth >= 2 else (
                axs.AXS_MARCH_PREFETCH | axs.AXS_MARCH_NT_STORE)
            lat.set_kernel("auto", depth=depth, flags=flags)
        except axs.AxsError as e:           # not available for this form
            log(f"[bench] {precision}: depth {depth} not applied ({e})")
    log(f"[bench] {precision}: {N}^3 initialised, {warmup} warm-up steps")
    if warmup:
        lat.step(warmup)
    log(f"[bench] {precision}: timing {steps} steps")
    t0 = time.perf_counter()
    lat.step(steps)                  # synchronous: returns after the stream drains
    wall = time.perf_counter() - t0
    kern_ms = lat.last_step_ms       # HIP-event average over the same steps
    scheme_b, alg_b = lat.bytes_per_site()
    L = N ** 3
    cfg = lat.kernel_config()
    res = {
        "kernel_config": cfg,
        "steps_per_launch": max(1, cfg["depth"]),
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": kern_ms,
        "site_updates_per_s": L * steps / wall,
        "alg_bytes_per_site": alg_b,
        "scheme_bytes_per_site": scheme_b,
        "init_s": {"first": init_first, "again": init_again},
    }
    log(f"[bench] {precision}: {res['ms_per_step']:.3f} ms/step wall, {kern_ms:.3f} ms kernel")
    return lat, res


def sampling_leg(np, lat, N, reps=3):
    """Time of each sampling call at a sampling point, as a run makes them
    (median over `reps` sampling points, one step apart, each call
    synchronous and in this order), with the effective HBM rate on the
    call's algorithmic bytes: strings read phi (FP32 8 B/site, FP64 16),
    background and energy read phi and phidot (16 / 32 B/site), the power
    spectrum is FFT-bound (rate on one read of phi2).  The coordinate list
    reuses the count's search of the same state.  In the FP32 form one pass
    over the stored (phi, u) state rebuilds phidot and sums both the
    background and the energy: the background call runs it, the energy call
    returns its kept result (no rate), and `background_plus_energy` is the
    cost of both at a sampling point in either form."""
    L = N ** 3
    f32 = lat.precision == "fp32"
    w = 4 if f32 else 8
    calls = {
        "string_counts": (lambda: lat.string_counts(), 2 * w),
        "strings_with_coords": (lambda: lat.strings(), 2 * w),
        "background": (lambda: lat.background(), 4 * w),
        "energy": (lambda: lat.energy(), 0 if f32 else 4 * w),
        "powerspec_log64": (lambda: lat.powerspec(binning="log", nbins=64), w),
        "powerspec_linear": (lambda: lat.powerspec(binning="linear"), w),
    }
    times = {name: [] for name in calls}
    for _ in range(reps):
        lat.step(1)               # a new state: nothing cached
        for name, (fn, _) in calls.items():
            t0 = time.perf_counter()
            fn()
            times[name].append(time.perf_counter() - t0)
    out = {}
    for name, (_, bps) in calls.items():
        t = sorted(times[name])[len(times[name]) // 2]
        out[name] = {"ms": 1e3 * t, "alg_GBs": (bps * L / t / 1e9) if bps else None}
    both = sorted(a + b for a, b in zip(times["background"], times["energy"]))
    out["background_plus_energy"] = {"ms": 1e3 * both[len(both) // 2]}
    return out


def _rel_l2_chunked(np, a, b, chunk=1 << 26):
    num = den = 0.0
    for lo in range(0, a.size, chunk):
        d = a[lo:lo + chunk] - b[lo:lo + chunk]
        num += float(np.dot(d, d))
        den += float(np.dot(b[lo:lo + chunk], b[lo:lo + chunk]))
    return num, den


def _pk_rel_err(np, a, b):
    """largest relative difference of a bin holding at least 1e-8 of the
    peak power (below that a bin holds rounding noise: the initial k^-1
    field has no power above pi/(4 dx))"""
    sel = b >= 1e-8 * b.max()
    return float(np.max(np.abs(a[sel] - b[sel]) / b[sel])) if sel.any() else 0.0


def config3_leg(np, lats, steps=1000, every=50, nbins=64):
    """The 3D configuration as a whole run in each form (the contexts of
    the timed legs, re-initialised): the device k^-1 field of seed 1998,
    tau = 1, `steps` steps with the string count and the log-binned power
    spectrum sampled every `every` steps before the step (s_run.cpp:83-85);
    the sampling path's one-time set-up (FFT plans) runs before the clock.
    Then the FP32 run is compared with the FP64 run: field rel-L2 and string
    count after the last step (the accuracy gate), energy, and the largest
    differences over the in-loop samples."""
    out, samples = {}, {}
    for prec, lat in lats.items():
        lat.init("kinv", seed=1998)
        lat.tau = 1.0
        lat.string_counts()
        lat.powerspec(binning="log", nbins=nbins)
        lat.init("kinv", seed=1998)
        lat.tau = 1.0
        counts, pks = [], []
        step_s = samp_s = 0.0
        t0 = time.perf_counter()
        for _ in range(0, steps, every):
            ts = time.perf_counter()
            counts.append(int(lat.string_counts()[0][0]))
            pks.append(lat.powerspec(binning="log", nbins=nbins)[1])
            tm = time.perf_counter()
            lat.step(every)
            step_s += time.perf_counter() - tm
            samp_s += tm - ts
        total = time.perf_counter() - t0
        samples[prec] = (counts, pks)
        L = lat.length()
        out[prec] = {"run_s": total, "step_s": step_s, "sample_s": samp_s,
                     "site_updates_per_s": L * steps / total,
                     "step_site_updates_per_s": L * steps / step_s,
                     "final_string_count": int(lat.string_counts()[0][0]),
                     "energy": [float(x) for x in lat.energy()]}
        log(f"[bench] config3 {prec}: {steps} steps + {len(counts)} samples in {total:.2f} s "
            f"(steps {step_s:.2f} s), {out[prec]['final_string_count']} strings")
    if set(lats) != {"fp32", "fp64"}:
        return out
    num = den = 0.0
    for comp in ("phi1", "phi2"):
        a = lats["fp32"].get_state(fields=(comp,))[comp]
        b = lats["fp64"].get_state(fields=(comp,))[comp]
        n_, d_ = _rel_l2_chunked(np, a, b)
        num, den = num + n_, den + d_
        del a, b
    rel_l2 = float(np.sqrt(num / den))
    n32, n64 = out["fp32"]["final_string_count"], out["fp64"]["final_string_count"]
    e32, e64 = sum(out["fp32"]["energy"]), sum(out["fp64"]["energy"])
    c32, p32 = samples["fp32"]
    c64, p64 = samples["fp64"]
    cnt_err = max(abs(a - b) / max(b, 1) for a, b in zip(c32, c64))
    pk_max = pk_pow = 0.0
    for a, b in zip(p32, p64):
        pk_max = max(pk_max, _pk_rel_err(np, a, b))
        pk_pow = max(pk_pow, float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))))
    count_rel = abs(n32 - n64) / max(n64, 1)
    out["accuracy_fp32_vs_fp64"] = {
        "steps": steps,
        "field_rel_l2": rel_l2,
        "string_count_fp32": n32,
        "string_count_fp64": n64,
        "string_count_rel_diff": count_rel,
        "energy_rel_diff": abs(e32 - e64) / abs(e64),
        "samples": len(c32),
        "sample_string_count_rel_diff_max": cnt_err,
        "pk_log64_max_rel_err": pk_max,
        "pk_log64_max_rel_err_floor": "bins holding >= 1e-8 of the sample's peak power",
        "pk_log64_power_weighted_err_max": pk_pow,
        "gate": {"field_rel_l2_max": 1e-4, "string_count_rel_diff_max": 0.01,
                 "pass": bool(rel_l2 <= 1e-4 and count_rel <= 0.01)},
    }
    out["fp32_speedup_over_fp64_run"] = out["fp64"]["run_s"] / out["fp32"]["run_s"]
    return out


ENSEMBLES = {
    # 512 2D lattices: 128 each of 128^2, 256^2, 512^2, 1024^2; 2000 steps,
    # strings and spectra of every realization every 500 steps
    "ensemble_2d": (2, [(128, 128), (256, 128), (512, 128), (1024, 128)], 2000, 500),
    # 48 3D lattices: 16 each of 128^3, 192^3, 256^3; 1500 steps, sampled
    # every 250 steps
    "ensemble_3d": (3, [(128, 16), (192, 16), (256, 16)], 1500, 250),
}


def ensemble_leg(np, axs, name, T0, steps_override=0, autotune=True):
    """All members advance together (one persistent launch per pass of
    several steps: the 3D and 2D multi-step kernels' work lists), launch
    shape from the autotuner; in-loop sampling of every realization (string
    count, 64-bin log power spectrum: one call each for the whole ensemble)
    at the stated cadence, the sampling path's one-time set-up (FFT plans)
    timed apart as sample_setup_s.  The FP32 form
    is timed, then per-realization string counts and spectra are compared
    with the FP64 form after the same steps."""
    nd, groups, steps, every = ENSEMBLES[name]
    if steps_override:
        steps = steps_override
        every = max(1, steps // 4)
    Ns = [N for N, cnt in groups for _ in range(cnt)]
    sites = sum(N ** nd for N in Ns)
    out = {"members": len(Ns), "sides": [N for N, _ in groups], "sites": sites, "steps": steps,
           "sample_every": every}
    counts, spectra = {}, {}
    for prec in ("fp32", "fp64"):
        lat = axs.Lattice(nd, members=Ns, precision=prec, T_initial=T0)
        for r in range(len(Ns)):
            lat.init("kinv", seed=1998 + r, member=r)
        lat.tau = 1.0
        if autotune:
            cfg, tuned_ms = lat.autotune(trials=3)
            log(f"[bench] {name} {prec}: {len(Ns)} members, tuned {cfg} ({tuned_ms:.3f} ms)")
        else:
            cfg = lat.kernel_config()
        # the sampling path's one-time set-up (rocFFT builds and plans its
        # kernels for each side on first use): one sample of the initial
        # state, timed on its own
        t0 = time.perf_counter()
        lat.string_counts()
        lat.powerspec_all(binning="log", nbins=64)
        setup_s = time.perf_counter() - t0
        step_s = sample_s = 0.0
        kern_ms_sum = 0.0
        done = nsamples = 0
        while done < steps:
            k = min(every, steps - done)
            t0 = time.perf_counter()
            lat.step(k)
            step_s += time.perf_counter() - t0
            kern_ms_sum += lat.last_step_ms * k
            done += k
            t0 = time.perf_counter()
            c = lat.string_counts()[0]
            pks = [x[1] for x in lat.powerspec_all(binning="log", nbins=64)]
            sample_s += time.perf_counter() - t0
            nsamples += 1
        counts[prec], spectra[prec] = c, pks
        out[prec] = {"site_updates_per_s": sites * steps / step_s, "ms_per_step": 1e3 * step_s / steps,
                     "kernel_ms": kern_ms_sum / steps, "config": cfg, "samples": nsamples,
                     "sample_s": sample_s, "sample_setup_s": setup_s}
        lat.close()
        log(f"[bench] {name} {prec}: {1e3 * step_s / steps:.3f} ms/step, "
            f"{nsamples} in-loop samples in {sample_s:.2f} s")
    c32 = counts["fp32"].astype(np.float64)
    c64 = counts["fp64"].astype(np.float64)
    rel = np.abs(c32 - c64) / np.maximum(c64, 1.0)
    out["string_count_rel_diff_max"] = float(rel.max())
    out["string_count_rel_diff_mean"] = float(rel.mean())
    # where the largest relative difference is (late in the run a small
    # lattice holds a handful of strings, so one string is a large fraction)
    # and the largest difference in strings
    w = int(np.argmax(rel))
    out["string_count_rel_diff_max_at"] = {"member": w, "side": Ns[w], "fp32": int(c32[w]),
                                           "fp64": int(c64[w])}
    out["string_count_abs_diff_max"] = int(np.abs(c32 - c64).max())
    out["string_count_members_equal"] = int(np.sum(c32 == c64))
    out["string_count_total_fp32"] = int(c32.sum())
    out["string_count_total_fp64"] = int(c64.sum())
    pk_err = pk_pow = 0.0
    for a, b in zip(spectra["fp32"], spectra["fp64"]):
        if np.any(b != 0):
            pk_err = max(pk_err, _pk_rel_err(np, a, b))
            pk_pow = max(pk_pow, float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))))
    out["pk_log64_max_rel_err"] = pk_err
    out["pk_log64_max_rel_err_floor"] = "bins holding >= 1e-8 of the member's peak power"
    out["pk_log64_power_weighted_err_max"] = pk_pow
    return out


# the reference's alternative step paths (FP64): the Hamiltonian form on the
# march kernel, the field-rescaled step with the spectral Laplacian (FFT-bound)
FORMS = {
    "hamiltonian_fp64": (dict(eom="hamiltonian"), 1024, 50),
    "spectral_laplacian_fp64": (dict(laplacian="spectral"), 512, 10),
}


def forms_leg(np, axs, T0):
    out = {}
    for name, (kw, N, steps) in FORMS.items():
        lat = axs.Lattice(3, N, precision="fp64", T_initial=T0, **kw)
        lat.init("kinv", seed=1998)
        lat.tau = 1.0
        lat.step(2)
        t0 = time.perf_counter()
        lat.step(steps)
        wall = time.perf_counter() - t0
        out[name] = {"N": N, "steps": steps, "ms_per_step": 1e3 * wall / steps,
                     "site_updates_per_s": N ** 3 * steps / wall,
                     "kernel_ms": lat.last_step_ms, "config": lat.kernel_config()}
        lat.close()
        log(f"[bench] {name}: {N}^3, {1e3 * wall / steps:.3f} ms/step")
    return out


def stencils_leg(np, axs, N, T0, steps=10):
    """The higher-order Laplacians (stencil_setting 1-3: radius 2-4) on the
    3D N^3 lattice in both forms: the library's default multi-step
    configuration against the one-step site kernel."""
    out = {}
    for prec in ("fp32", "fp64"):
        for st in (1, 2, 3):
            lat = axs.Lattice(3, N, precision=prec, T_initial=T0, stencil_setting=st)
            lat.init("kinv", seed=1998)
            lat.tau = 1.0
            cfg = lat.kernel_config()
            lat.step(2 * cfg["depth"])
            t0 = time.perf_counter()
            lat.step(steps)
            wall = time.perf_counter() - t0
            kern = lat.last_step_ms
            lat.set_kernel("simple")
            lat.step(1)
            t1 = time.perf_counter()
            lat.step(3)
            one = (time.perf_counter() - t1) / 3
            out[f"{prec}_stencil{st}"] = {
                "N": N, "steps": steps, "ms_per_step": 1e3 * wall / steps,
                "site_updates_per_s": N ** 3 * steps / wall, "kernel_ms": kern,
                "config": cfg, "one_step_kernel_ms_per_step": 1e3 * one,
                "speedup_over_one_step": one / (wall / steps)}
            lat.close()
            log(f"[bench] stencil {st} {prec}: {N}^3, {1e3 * wall / steps:.3f} ms/step "
                f"(one-step kernel {1e3 * one:.3f})")
    return out


# The reference's own small configurations as whole runs (FP64, field from
# the host k^-1 generator, sampling at the stated cadence): 2D 256^2, 500
# steps, strings every 50; 3D 256^3, 1000 steps, strings and energy every 20
WORKFLOWS = {
    "config0_2d_256": (2, 256, 500, 50, False),
    "config1_3d_256": (3, 256, 1000, 20, True),
}


def workflow_leg(np, axs, T0, cpu_rate_3d=None):
    """Wall time of each configuration run on the GPU (steps plus the
    in-loop sampling, as the driver makes the calls): configs 0 and 1 in
    the FP64 form from the host k^-1 generator's field; config 0 also on
    the CPU oracle, config 1 on the CPU estimated from the measured
    reference-structured step rate (steps only, no sampling)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {}
    for name, (nd, N, steps, every, energy) in WORKFLOWS.items():
        p1, p2, _, _ = axs.host_init("kinv", nd, N, T_initial=T0, seed=1998)
        z = np.zeros(N ** nd)
        lat = axs.Lattice(nd, N, precision="fp64", T_initial=T0)
        lat.set_state(p1, p2, z, z, tau=1.0)
        lat.step(1)                                   # code objects, first touch
        lat.set_state(p1, p2, z, z, tau=1.0)
        t0 = time.perf_counter()
        nstr = 0
        for done in range(0, steps, every):
            nstr = lat.strings(coords=False)[0]
            if energy:
                lat.energy()
            lat.step(every)
        gpu_s = time.perf_counter() - t0
        lat.close()
        rec = {"ndims": nd, "N": N, "steps": steps, "sample_every": every,
               "gpu_s": gpu_s, "gpu_site_updates_per_s": N ** nd * steps / gpu_s,
               "last_string_count": int(nstr)}
        if nd == 2:
            st = O.State(p1, p2, z, z)
            pp, A = O.params(2, N, T_initial=T0), O.Matrix(2, N)
            t0 = time.perf_counter()
            for done in range(0, steps, every):
                O.cores(O.axion(st.phi1, st.phi2), 2, N)
                st.run(pp, A, every)
            rec["cpu_s"] = time.perf_counter() - t0
            rec["cpu_kind"] = "oracle, reference structure, measured"
        elif cpu_rate_3d:
            rec["cpu_s"] = N ** nd * steps / cpu_rate_3d
            rec["cpu_kind"] = "oracle step rate at 256^3 x steps (sampling not included)"
        if "cpu_s" in rec:
            rec["speedup"] = rec["cpu_s"] / gpu_s
        out[name] = rec
        log(f"[bench] {name}: GPU {gpu_s:.3f} s" +
            (f", CPU {rec['cpu_s']:.2f} s" if "cpu_s" in rec else ""))
    return out


def snapshot_leg(np, axs, N, T0, steps=200):
    """Field snapshot of the 1024^3 FP32 run (phi1, phi2 of every site,
    8.6 GB each as float64 on the host): the synchronous read-back against
    the asynchronous one (device staging copy, transfer on a second stream,
    unpack on a host thread) overlapped with `steps` steps."""
    import threading
    lat = axs.Lattice(3, N, precision="fp32", T_initial=T0)
    lat.init("kinv", seed=1998)
    lat.tau = 1.0
    lat.step(8)
    t0 = time.perf_counter()
    lat.step(steps)
    plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat.get_state(fields=("phi1", "phi2"))
    sync_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat.snapshot_begin(fields=("phi1", "phi2"))
    box = {}
    th = threading.Thread(target=lambda: box.update(r=lat.snapshot_end()))
    th.start()
    lat.step(steps)
    steps_done = time.perf_counter() - t0
    th.join()
    total = time.perf_counter() - t0
    lat.close()
    del box
    return {"N": N, "fields": "phi1,phi2", "steps": steps,
            "steps_alone_s": plain, "sync_read_back_s": sync_s,
            "async_steps_s": steps_done, "async_total_s": total,
            "hidden_fraction": max(0.0, min(1.0, (plain + sync_s - total) / sync_s))}


def cpu_baseline(np, seconds: float):
    """The CPU oracle on the host's threads, on bounded samples of the
    workload: the reference-structured FP64 step (five sweeps + CSR SpMV,
    the reference's own structure) at 3D 256^3 -- the reported value -- and
    a fused two-sweep stencil stepper (same bits) at 2D 256^2, 3D 256^3 and
    3D 512^3, the last standing in for 1024^3 per site (its FP64 state and
    the reference's int32 COO matrix do not fit the reference structure)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import axs
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)

    def timed(run, budget, sites):
        run(1)                                  # first touch
        n, t = 0, 0.0
        while t < budget and n < 5000:
            t0 = time.perf_counter()
            run(1)
            t += time.perf_counter() - t0
            n += 1
        return sites * n / t, n, t

    Ns = 256
    p1, p2, _, _ = axs.host_init("kinv", 3, Ns, seed=1998)
    z = np.zeros(Ns ** 3)
    p = O.params(3, Ns)
    A = O.Matrix(3, Ns)
    st = O.State(p1, p2, z, z)
    rate, n, t = timed(lambda k: st.run(p, A, k), 0.5 * seconds, Ns ** 3)
    del A
    sf = O.State(p1, p2, z, z)
    f256, _, _ = timed(lambda k: sf.run_fused(p, k), 0.15 * seconds, Ns ** 3)
    q1, q2, _, _ = axs.host_init("kinv", 2, 256, seed=1998)
    z2 = np.zeros(256 ** 2)
    p2d = O.params(2, 256)
    A2 = O.Matrix(2, 256)
    s2a, s2b = O.State(q1, q2, z2, z2), O.State(q1, q2, z2, z2)
    r2d, _, _ = timed(lambda k: s2a.run(p2d, A2, 20 * k), 0.05 * seconds, 20 * 256 ** 2)
    f2d, _, _ = timed(lambda k: s2b.run_fused(p2d, 20 * k), 0.05 * seconds, 20 * 256 ** 2)
    del sf, s2a, s2b, A2
    N5 = 512
    rng = np.random.default_rng(1998)
    z5 = np.zeros(N5 ** 3)
    s5 = O.State(rng.standard_normal(N5 ** 3), rng.standard_normal(N5 ** 3), z5, z5)
    f512, _, _ = timed(lambda k: s5.run_fused(O.params(3, N5), k), 0.2 * seconds, N5 ** 3)
    del s5
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": rate,
        "unit": "site-updates/s",
        "cores": threads,
        "host": {"cpu_model": model, "logical_cpus": os.cpu_count()},
        "kind": "port",
        "sample": f"3D {Ns}^3 FP64, {n} reference-structured steps (drift, 2x CSR SpMV "
                  f"Laplacian, potential, kick, copy) of the CPU oracle, {t:.1f} s, "
                  f"OMP_NUM_THREADS={threads}",
        "reference_structured_2d_256": r2d,
        "fused_stencil": {"2d_256": f2d, "3d_256": f256, "3d_512": f512,
                          "note": "two-sweep fused stepper, bit-identical to the reference "
                                  "structure; 3d_512 stands in for 1024^3 per site"},
    }


def copy_calibration(nbytes: int) -> int:
    """One device-to-device copy of nbytes through the HIP runtime: a copy
    kernel that reads and writes exactly nbytes, to scale FETCH_SIZE and
    WRITE_SIZE readings against."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            print("hipMalloc failed", file=sys.stderr)
            return 1
    hip.hipMemset(a, 0, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "rc": rc}), flush=True)
    return 0 if rc == 0 else 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--T-initial", type=float, default=4.0)
    ap.add_argument("--skip-fp64", action="store_true", help="time the FP32 form only")
    ap.add_argument("--forms", default="fp32,fp64",
                    help="forms to time (fp32 is the headline; fp64 alone for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (for calibrating HBM "
                         "traffic counters) and exit")
    ap.add_argument("--skip-sampling", action="store_true",
                    help="skip timing the sampling calls (strings, reductions, spectra)")
    ap.add_argument("--skip-snapshot", action="store_true",
                    help="skip the asynchronous-snapshot overlap leg")
    ap.add_argument("--skip-forms", action="store_true",
                    help="skip the Hamiltonian-form and spectral-Laplacian legs")
    ap.add_argument("--skip-workflows", action="store_true",
                    help="skip the whole-run timings of the 256^2 and 256^3 configurations")
    ap.add_argument("--skip-config3", action="store_true",
                    help="skip the whole 1000-step run of both forms and its accuracy gate")
    ap.add_argument("--config3-steps", type=int, default=1000,
                    help="steps of the config-3 run (the accuracy gate is stated at 1000)")
    ap.add_argument("--skip-stencils", action="store_true",
                    help="skip the higher-order Laplacian legs")
    ap.add_argument("--skip-ensembles", action="store_true",
                    help="skip the 2D (512-member) and 3D (48-member) ensemble legs")
    ap.add_argument("--no-ensemble-autotune", action="store_true",
                    help="run the ensembles in the library's default launch shape")
    ap.add_argument("--ensemble-steps", type=int, default=0,
                    help="override the ensembles' step counts (2000 / 1500)")
    ap.add_argument("--depth", type=int, default=0,
                    help="time steps per step-kernel launch (0: the library's default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design only)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    import numpy as np
    import axs

    N, K, W = args.N, args.steps, args.warmup
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    if forms == ["fp64"]:
        lat64, r64 = gpu_leg(np, axs, N, "fp64", W, K, args.T_initial, args.depth)
        lat64.close()
        print(json.dumps({"form": "fp64", **r64}), flush=True)
        return 0
    lat32, r32 = gpu_leg(np, axs, N, "fp32", W, K, args.T_initial, args.depth)
    out64 = None
    config3 = None
    sampling = None
    lat64 = None
    if not args.skip_fp64 and "fp64" in forms:
        lat64, r64 = gpu_leg(np, axs, N, "fp64", W, K, args.T_initial, args.depth)
        launch64 = r64["kernel_ms"] * r64["steps_per_launch"]
        out64 = {
            "value": r64["site_updates_per_s"],
            "ms_per_step": r64["ms_per_step"],
            "kernel_ms": r64["kernel_ms"],
            "steps_per_launch": r64["steps_per_launch"],
            "launch_ms": launch64,
            # compulsory bytes of a pass (phi, phidot, ker read and written
            # once: 96 B per site) over the launch time
            "roofline_frac": PASS_BYTES["fp64"] * N ** 3 / (launch64 * 1e-3)
            / (HBM_PEAK_GBS * 1e9),
            "scheme_bytes_per_site": r64["scheme_bytes_per_site"],
            "kernel_config": r64["kernel_config"],
            "init_s": r64["init_s"],
        }
        if not args.skip_sampling:
            try:
                sampling = {"fp32": sampling_leg(np, lat32, N),
                            "fp64": sampling_leg(np, lat64, N)}
                log("[bench] sampling: " + ", ".join(
                    f"{k} {v['ms']:.1f} ms" for k, v in sampling["fp32"].items()))
            except Exception as e:          # reported, the headline line still prints
                sampling = {"error": str(e)}
    if not args.skip_config3:
        try:
            lats = {"fp32": lat32} if lat64 is None else {"fp32": lat32, "fp64": lat64}
            config3 = config3_leg(np, lats, steps=args.config3_steps)
            a = config3.get("accuracy_fp32_vs_fp64")
            if a:
                log(f"[bench] config3 accuracy after {a['steps']} steps: rel-L2 "
                    f"{a['field_rel_l2']:.3g}, strings {a['string_count_fp32']} vs "
                    f"{a['string_count_fp64']}, gate {'pass' if a['gate']['pass'] else 'FAIL'}")
        except Exception as e:              # reported, the headline line still prints
            config3 = {"error": str(e)}
    if lat64 is not None:
        lat64.close()
    lat32.close()

    L = N ** 3
    spl = r32["steps_per_launch"]
    launch_ms = r32["kernel_ms"] * spl
    achieved = PASS_BYTES["fp32"] * L / (launch_ms * 1e-3) / 1e9
    update_rate = r32["alg_bytes_per_site"] * L * spl / (launch_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tfile = TRAFFIC_FILES.get(spl)
    if tfile and os.path.exists(tfile):
        try:
            t = json.load(open(tfile))
            same = (int(t.get("N", -1)) == N and int(t.get("steps_per_launch", 1)) == spl and
                    t.get("kernel_config") == r32["kernel_config"])
            traffic_src = {"file": os.path.relpath(tfile, ROOT),
                           "kernel": t.get("kernel"), "kernel_config": t.get("kernel_config"),
                           "matches_run": bool(same)}
            if same:
                traffic = t.get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None

    ens = None
    if not args.skip_ensembles:
        ens = {}
        for name in ENSEMBLES:
            try:
                ens[name] = ensemble_leg(np, axs, name, args.T_initial, args.ensemble_steps,
                                         not args.no_ensemble_autotune)
            except Exception as e:          # reported, the headline line still prints
                ens[name] = {"error": str(e)}

    forms_out = None
    if not args.skip_forms:
        try:
            forms_out = forms_leg(np, axs, args.T_initial)
        except Exception as e:              # reported, the headline line still prints
            forms_out = {"error": str(e)}

    stencils = None
    if not args.skip_stencils:
        try:
            stencils = stencils_leg(np, axs, N, args.T_initial)
        except Exception as e:              # reported, the headline line still prints
            stencils = {"error": str(e)}

    snap = None
    if not args.skip_snapshot:
        try:
            snap = snapshot_leg(np, axs, N, args.T_initial)
            log(f"[bench] snapshot: sync {snap['sync_read_back_s']:.2f} s, async total "
                f"{snap['async_total_s']:.2f} s for {snap['steps']} steps + snapshot")
        except Exception as e:
            snap = {"error": str(e)}

    cpu = None
    if not args.skip_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(np, args.cpu_seconds)

    workflows = None
    if not args.skip_workflows:
        try:
            workflows = workflow_leg(np, axs, args.T_initial, cpu["value"] if cpu else None)
        except Exception as e:              # reported, the headline line still prints
            workflows = {"error": str(e)}

    line = {
        "metric": "lattice site-updates/s (3D 1024^3, FP32 form; FP64 form alongside)",
        "value": r32["site_updates_per_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: k^-1 Gaussian random field, seed 1998, generated on the device",
        "config": {
            "workload": f"3D {N}^3 periodic lattice velocity-Verlet step, FP32 phi/velocity, "
                        "lambda=1, dx=0.5, dt=0.1, T_initial=%g, 7-point Laplacian" % args.T_initial,
            "N": N, "ndims": 3, "precision": "fp32", "T_initial": args.T_initial,
            "stencil_setting": 0, "init": "kinv seed 1998",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "basis": "compulsory bytes of one launch: 32 B per site (phi1, phi2, u1, u2 "
                     "read and written once in float) x N^3, over the HIP-event average "
                     "launch duration; a launch advances steps_per_launch steps, so at "
                     "the same time per step a 5-step pass reads lower than a 4-step one",
            "steps_per_launch": spl,
            "launch_ms": launch_ms,
            "site_update_rate": update_rate,
            "traffic_source": traffic_src,
            "traffic_rate": (traffic / (launch_ms * 1e-3) / 1e9) if traffic else None,
            "traffic_frac": (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
        },
        "kernel_config": r32["kernel_config"],
        "kernel_ms_per_step": r32["kernel_ms"],
        "init_s": r32["init_s"],
        "fp64": out64,
        "fp32_speedup_over_fp64": (r32["site_updates_per_s"] / out64["value"]) if out64 else None,
        "accuracy_fp32_vs_fp64": (config3 or {}).get("accuracy_fp32_vs_fp64"),
        "config3": config3,
        "sampling": sampling,
        "ensembles": ens,
        "forms": forms_out,
        "stencils": stencils,
        "snapshot": snap,
        "workflows": workflows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

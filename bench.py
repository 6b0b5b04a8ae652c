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
  roofline     : the step kernel's achieved rate: algorithmic bytes per
                 launch (32 B per site-update in FP32: read + write of phi1,
                 phi2 and the two velocities, times the site-updates of one
                 launch: N^3 per time step it advances) over the launch's
                 average duration, measured with HIP events on the library's
                 stream.  The default FP32 kernel advances four time steps
                 per launch, reading and writing the state once, so this
                 figure exceeds the HBM peak; `traffic` (measured HBM bytes
                 per launch) and `traffic_rate` give the bytes actually
                 moved;
  cpu_baseline : the CPU oracle (reference-structured FP64 step: drift,
                 two CSR sparse-matrix Laplacians, potential, kick, copy, with
                 OpenMP) timed on this host over a bounded sample;
  accuracy     : FP32 vs FP64 field rel-L2, string count, energy and
                 power-spectrum differences after the same W+K steps.
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
# steps per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)
TRAFFIC_FILES = {1: os.path.join(ROOT, "profiles", "traffic_march_fp32_3d_1024.json"),
                 4: os.path.join(ROOT, "profiles", "traffic_multi_step_fp32_3d_1024.json")}


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
            lat.set_kernel("auto", depth=depth)
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


def accuracy(np, lat32, lat64):
    s32 = lat32.get_state(fields=("phi1", "phi2"))
    s64 = lat64.get_state(fields=("phi1", "phi2"))
    num = np.sqrt(np.sum((s32["phi1"] - s64["phi1"]) ** 2) + np.sum((s32["phi2"] - s64["phi2"]) ** 2))
    den = np.sqrt(np.sum(s64["phi1"] ** 2) + np.sum(s64["phi2"] ** 2))
    del s32, s64
    n32 = lat32.strings(coords=False)[0]
    n64 = lat64.strings(coords=False)[0]
    e32, e64 = lat32.energy(), lat64.energy()
    _, pk32, c32 = lat32.powerspec(binning="log", nbins=64)
    _, pk64, c64 = lat64.powerspec(binning="log", nbins=64)
    sel = c64 > 0
    pk_err = float(np.max(np.abs(pk32[sel] - pk64[sel]) / np.abs(pk64[sel])))
    return {
        "field_rel_l2": float(num / den),
        "string_count_fp32": int(n32),
        "string_count_fp64": int(n64),
        "string_count_rel_diff": float(abs(n32 - n64) / max(n64, 1)),
        "energy_rel_diff": float(abs(e32.sum() - e64.sum()) / abs(e64.sum())),
        "pk_log64_max_rel_err": pk_err,
    }


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
    out["string_count_total_fp32"] = int(c32.sum())
    out["string_count_total_fp64"] = int(c64.sum())
    pk_err = 0.0
    for a, b in zip(spectra["fp32"], spectra["fp64"]):
        sel = b != 0
        if sel.any():
            pk_err = max(pk_err, float(np.max(np.abs(a[sel] - b[sel]) / np.abs(b[sel]))))
    out["pk_log64_max_rel_err"] = pk_err
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
    the FP64 form, config 3 in the FP32 form; config 0 also on the CPU
    oracle, configs 1 and 3 on the CPU estimated from the measured
    reference-structured step rate (steps only, no sampling)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from fields import kinv_field
    out = {}
    for name, (nd, N, steps, every, energy) in WORKFLOWS.items():
        p1, p2 = kinv_field(nd, N, seed=1998)
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
    # config 3 as a whole run in the FP32 form: 3D 1024^3, 1000 steps, the
    # string count and the 64-bin log spectrum every 50 steps, from the
    # device k^-1 field (the host generator would take minutes at 1024^3);
    # the sampling path's one-time set-up (FFT plans) runs before the clock
    N, steps, every = 1024, 1000, 50
    lat = axs.Lattice(3, N, precision="fp32", T_initial=T0)
    lat.init("kinv", seed=1998)
    lat.string_counts()
    lat.powerspec(binning="log", nbins=64)
    lat.init("kinv", seed=1998)
    lat.tau = 1.0
    t0 = time.perf_counter()
    nstr = 0
    for done in range(0, steps, every):
        nstr = int(lat.string_counts()[0][0])
        lat.powerspec(binning="log", nbins=64)
        lat.step(every)
    gpu_s = time.perf_counter() - t0
    lat.close()
    rec = {"ndims": 3, "N": N, "steps": steps, "sample_every": every, "precision": "fp32",
           "samples": "string count + 64-bin log power spectrum", "gpu_s": gpu_s,
           "gpu_site_updates_per_s": N ** 3 * steps / gpu_s, "last_string_count": nstr}
    if cpu_rate_3d:
        rec["cpu_s"] = N ** 3 * steps / cpu_rate_3d
        rec["cpu_kind"] = "oracle FP64 step rate at 256^3 x steps (sampling not included)"
        rec["speedup"] = rec["cpu_s"] / gpu_s
    out["config3_3d_1024_fp32"] = rec
    log(f"[bench] config3_3d_1024_fp32: GPU {gpu_s:.3f} s")
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
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from fields import kinv_field
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
    p1, p2 = kinv_field(3, Ns, seed=1998)
    z = np.zeros(Ns ** 3)
    p = O.params(3, Ns)
    A = O.Matrix(3, Ns)
    st = O.State(p1, p2, z, z)
    rate, n, t = timed(lambda k: st.run(p, A, k), 0.5 * seconds, Ns ** 3)
    del A
    sf = O.State(p1, p2, z, z)
    f256, _, _ = timed(lambda k: sf.run_fused(p, k), 0.15 * seconds, Ns ** 3)
    q1, q2 = kinv_field(2, 256, seed=1998)
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
    ap.add_argument("--warmup", type=int, default=10)
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
    acc = None
    sampling = None
    if not args.skip_fp64 and "fp64" in forms:
        lat64, r64 = gpu_leg(np, axs, N, "fp64", W, K, args.T_initial, args.depth)
        try:
            log("[bench] comparing FP32 with FP64")
            acc = accuracy(np, lat32, lat64)
        except Exception as e:              # reported, the headline line still prints
            acc = {"error": str(e)}
        if not args.skip_sampling:
            try:
                sampling = {"fp32": sampling_leg(np, lat32, N),
                            "fp64": sampling_leg(np, lat64, N)}
                log("[bench] sampling: " + ", ".join(
                    f"{k} {v['ms']:.1f} ms" for k, v in sampling["fp32"].items()))
            except Exception as e:          # reported, the headline line still prints
                sampling = {"error": str(e)}
        lat64.close()
        out64 = {
            "value": r64["site_updates_per_s"],
            "ms_per_step": r64["ms_per_step"],
            "kernel_ms": r64["kernel_ms"],
            "roofline_frac_alg": (r64["alg_bytes_per_site"] * N ** 3 / (r64["kernel_ms"] * 1e-3))
            / (HBM_PEAK_GBS * 1e9),
            "scheme_bytes_per_site": r64["scheme_bytes_per_site"],
            "kernel_config": r64["kernel_config"],
            "init_s": r64["init_s"],
        }
    lat32.close()

    L = N ** 3
    spl = r32["steps_per_launch"]
    launch_ms = r32["kernel_ms"] * spl
    achieved = r32["alg_bytes_per_site"] * L * spl / (launch_ms * 1e-3) / 1e9
    traffic = None
    tfile = TRAFFIC_FILES.get(spl)
    if tfile and os.path.exists(tfile):
        try:
            t = json.load(open(tfile))
            if int(t.get("N", -1)) == N and int(t.get("steps_per_launch", 1)) == spl:
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
            "steps_per_launch": spl,
            "launch_ms": launch_ms,
            "traffic_rate": (traffic / (launch_ms * 1e-3) / 1e9) if traffic else None,
            "traffic_frac": (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
        },
        "kernel_config": r32["kernel_config"],
        "kernel_ms_per_step": r32["kernel_ms"],
        "init_s": r32["init_s"],
        "fp64": out64,
        "fp32_speedup_over_fp64": (r32["site_updates_per_s"] / out64["value"]) if out64 else None,
        "accuracy_fp32_vs_fp64": acc,
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

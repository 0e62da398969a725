"""Benchmark: timesteps/s of the 256^3 FP64 fractional step on one MI355X.

Workload (the headline configuration): FP64, 256^3 on [0, 2 pi)^3, random
solenoidal IC (E(k) ~ k^4 exp(-2 (k/4)^2), u_rms = 1, seed 2024), nu = 1/800,
dt = 2e-3.  One "step" is one iteration of exec/ns.cpp:167-217 on the fused
kernel schedule.  Inputs are resident in HBM when the timed region starts.

`value` is the CHOSEN FORM: the start is real, so the solver stores and steps
only the planes kz < nz/2 of the Hermitian spectrum (an R2C-equivalent
storage; the same velocity to rounding), and the tuned kernel E keeps the
merged AB2 history Y = (V1 + lap1) - K/2 N0 in place of V1 and N0 and emits
five fields from which the next y pass forms the vorticity: 40 field passes,
320 N^3 algorithmic bytes per step (42 / 336 N^3 with six emitted fields,
48 / 384 N^3 with a kernel E that keeps V1 and N0).  The PLAIN method -- the full complex spectrum, the
reference's C2C arithmetic operation for operation -- is reported beside it
under "full_spectrum" (640 to 768 N^3 fused bytes per step against the
1056 N^3-byte contract of 9 3D FFTs + 12 array passes), with its own roofline
and kernel table.

The reference scheme (explicit AB2 advection, Nyquist zeroing only) overflows
at this dt between steps 60 and 70 (CPU oracle and GPU alike), so the default
window (5 warmup + 50 timed steps) stays finite; "stable_dt" repeats the run
at dt = 5e-4 over 200 steps.  Per-step cost does not depend on the values.

Prints ONE JSON line: the metric, a roofline object for the dominant kernel
(algorithmic bytes / average launch time from HIP events recorded on the
launch stream in the timed region), the measured HBM copy ceiling, the CPU
baseline (the oracle's op-by-op restatement on the host cores, virtual slab
ranks x threads swept, -march=native build), and extras: the FP32 form, the
64^3 LES step rates, the standalone 3D FFT GB/s at 256^3 and 512^3.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
NU, DT, SEED, K0 = 1.0 / 800.0, 2e-3, 2024, 4.0
# HBM bytes per launch at 256^3 measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (in separate runs), each scaled by the same counter's reading for
# a plain copy kernel of known size; the summary and the factors are in the
# file named by TRAFFIC_SOURCE.  Keyed by the storage layout.
PROF_EVERY = 5  # steps of the timed region whose launches carry HIP events: 1 in PROF_EVERY
TRAFFIC_SOURCE = "profiles/pmc_traffic_ns256.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"
MEASURED_TRAFFIC_256 = {
    "full": {"ns_cross_z": 2.416e9, "ns_fourier_x": 5.638e9, "ns_inv_x": 2.416e9,
             "ns_y_inv": 3.221e9, "ns_y_fwd": 1.611e9},
    "half": {"ns_cross_z": 1.208e9, "ns_fourier_x": 2.820e9, "ns_inv_x": 1.208e9,
             "ns_y_inv": 1.611e9, "ns_y_fwd": 0.805e9},
}
# kernel E with the merged AB2 history, keyed by (layout, form): form 1 emits
# six fields, form 2 five (the full layout not measured)
MEASURED_TRAFFIC_256_MERGED_E = {("full", 1): None, ("half", 1): 2.014e9, ("full", 2): None, ("half", 2): 1.880e9}
# algorithmic field passes of each kernel class (one complex element read or
# written per stored grid point = one pass: 16 B in FP64, 8 B in FP32)
PASSES = {
    "ns_inv_x": 3 + 6,        # read V, write 6 x-transformed fields
    "ns_y_inv": 6 + 6,        # y pass of 6 fields
    "ns_cross_z": 6 + 3,      # z inverse of 6, cross, z forward of 3
    "ns_y_fwd": 3 + 3,        # y pass of 3 fields
    "ns_fourier_x": 9 + 12,   # read X, V1, N0; write V1, N0 and next step's 6 x-inverse fields
}
# kernel E with the merged AB2 history: read X and Y, write Y and the 6 fields
# (form 1) or 5 fields (form 2: V and d/dx Vy, d/dx Vz); with form 2 the y
# inverse reads those 5 and writes V and the vorticity (6)
PASSES_MERGED_E = {1: 3 + 3 + 3 + 6, 2: 3 + 3 + 3 + 5}
PASSES_Y_CURL = 5 + 6
NS_KERNELS = ("ns_inv_x", "ns_y_inv", "ns_cross_z", "ns_y_fwd", "ns_fourier_x")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--fp32-steps", type=int, default=100)
    p.add_argument("--les-steps", type=int, default=200, help="0 skips the LES step-rate line")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fft", action="store_true")
    p.add_argument("--no-full", action="store_true", help="skip the full-spectrum FP64 line")
    p.add_argument("--no-copy", action="store_true", help="skip the HBM copy ceiling")
    p.add_argument("--cpu-steps", type=int, default=3)
    p.add_argument("--stable-dt", type=float, default=5e-4)
    p.add_argument("--stable-steps", type=int, default=200)
    p.add_argument("--no-tune", action="store_true", help="use the default schedules")
    return p.parse_args()


def prof_total(L, name):
    ms = ctypes.c_double()
    cnt = ctypes.c_long()
    L.sh_prof_query(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, cnt.value


def run_ns(S, n, precision, warmup, steps, profile, dt=DT, layout=1):
    import torch
    L = S.lib()
    V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_solenoidal_ic_z(S.ptr(V0), n, n, n, SEED, K0, 1.0, S.stream())
    ns = S.NS((n, n, n), NU, dt, precision, fused=True)
    L.sh_ns_set_layout(ns.h, layout)
    ns.set_initial(V0)
    del V0
    ns.startup()
    half = ctypes.c_int(0)
    L.sh_ns_get_layout(ns.h, ctypes.byref(half))
    if TUNE:
        L.sh_ns_autotune(ns.h, 3)
    form = ctypes.c_int(0)
    ns.steps(warmup)
    torch.cuda.synchronize()
    if profile:
        L.sh_prof_reset()
    t0 = time.perf_counter()
    if profile:
        # HIP events around every launch of one step in PROF_EVERY: each event
        # record costs the stream a few microseconds, so the sampled steps keep
        # the timed region within ~0.5% of an unprofiled one
        for i in range(steps):
            L.sh_prof_enable(1 if i % PROF_EVERY == 0 else 0)
            ns.steps(1)
    else:
        ns.steps(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.sh_ns_get_history_form(ns.h, ctypes.byref(form))
    kern = {}
    if profile:
        L.sh_prof_enable(0)
        for name in NS_KERNELS:
            kern[name] = prof_total(L, name)
    div = ns.div_average()
    V = ns.velocity()
    e = ctypes.c_double()
    L.sh_spectrum_z(S.ptr(V), n, n, n, NU, None, 0, ctypes.byref(e), None, S.stream())
    del V
    ns.close()
    torch.cuda.empty_cache()
    return t1 - t0, kern, {"div_avg": div, "energy": e.value, "half": bool(half.value),
                           "form": form.value}


def kernel_table(kern, n, half, form=0, prec_bytes=16):
    """per-kernel average launch time, algorithmic bytes and GB/s; the dominant
    kernel's roofline object (form: the solver's history form, 0 / 1 / 2)"""
    layout = "half" if half else "full"
    pts = float(n) ** 3 * (0.5 if half else 1.0)
    passes = dict(PASSES)
    if form:
        passes["ns_fourier_x"] = PASSES_MERGED_E[form]
    if form == 2:
        passes["ns_y_inv"] = PASSES_Y_CURL
    kernels = {}
    dom, dom_t = None, -1.0
    for name, (ms, cnt) in kern.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        byts = passes[name] * prec_bytes * pts
        kernels[name] = {"avg_ms": round(avg_ms, 5), "launches": cnt, "algo_GB": round(byts / 1e9, 4),
                         "GBps": round(byts / (avg_ms * 1e-3) / 1e9, 1)}
        if ms > dom_t:
            dom, dom_t = name, ms
    roof = None
    if dom:
        k = kernels[dom]
        traffic = MEASURED_TRAFFIC_256[layout].get(dom) if n == 256 else None
        if form and dom == "ns_fourier_x":
            traffic = MEASURED_TRAFFIC_256_MERGED_E[(layout, form)] if n == 256 else None
        roof = {"bound": "hbm", "kernel": dom, "achieved": k["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(k["GBps"] / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "traffic_source": TRAFFIC_SOURCE if traffic else None,
                "algo_bytes_per_launch": passes[dom] * prec_bytes * pts,
                "algo_bytes_formula": "%d field passes x %d B x %s" % (passes[dom], prec_bytes,
                                                                       "N^3/2" if half else "N^3"),
                "history_form": form if dom == "ns_fourier_x" else 0,
                "avg_launch_ms": k["avg_ms"], "timing": "HIP events on the launch stream around every launch of 1 in %d steps of the timed region"
                % PROF_EVERY}
    step_passes = sum(passes[k] for k, (ms, cnt) in kern.items() if cnt > 0)
    return kernels, roof, step_passes * prec_bytes * pts, step_passes


def fft_bench(S, n, reps):
    """exec/fft.cpp:52-58: min over repeats of forward() + inverse()."""
    import torch
    L = S.lib()
    if TUNE:
        L.sh_autotune_fft3d(n, n, n, S.SH_F64, 3, S.stream())
    x = torch.empty((n, n, n), dtype=torch.complex128, device="cuda")
    L.sh_prng_complex_z(S.ptr(x), x.numel(), 7, S.stream())
    best = 1e30
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.fft3d(x, -1, 0)
        S.fft3d(x, +1, S.SH_NORMALIZE)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    del x
    torch.cuda.empty_cache()
    per_fft = best / 2
    gbps = 96.0 * n ** 3 / per_fft / 1e9
    return per_fft, gbps


def copy_ceiling(S, gib=1, reps=5):
    """The project's streaming copy kernel (sh_stream_copy) over `gib` GiB:
    the best read+write GB/s over a few launch shapes, timed with HIP events
    on its stream -- the bandwidth a streaming kernel reaches on this GPU."""
    import torch
    L = S.lib()
    nb = gib << 30
    src = torch.empty(nb, dtype=torch.uint8, device="cuda")
    dst = torch.empty(nb, dtype=torch.uint8, device="cuda")
    L.sh_prng_complex_z(S.ptr(src), nb // 16, 5, S.stream())
    best = None
    for blocks in (0, 1024, 2048, 4096, 8192):
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            L.sh_stream_copy(S.ptr(dst), S.ptr(src), nb, blocks, S.stream())
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        gbps = 2.0 * nb / (min(ts) * 1e-3) / 1e9
        if best is None or gbps > best["GBps"]:
            best = {"GBps": round(gbps, 1), "frac_of_peak": round(gbps / HBM_PEAK_GBPS, 4),
                    "bytes_moved": 2 * nb, "blocks": blocks or "one trip per lane", "ms": round(min(ts), 4)}
    assert torch.equal(src, dst)
    del src, dst
    torch.cuda.empty_cache()
    best["kernel"] = "k_stream_copy (16 B per lane, nontemporal load/store)"
    return best


def cpu_baseline(n, steps):
    """The oracle's op-by-op step (exec/ns.cpp:132-217 restated) on the host:
    the real start-up, then `steps` timed steps at each virtual slab-rank count
    R in {1, the largest divisor of n <= threads} (src/Scalar.cpp:446-626:
    R rank threads on their slabs, the slab puts as copies; R = 1 runs every
    loop on all OpenMP threads).  Built with -march=native on this host when
    the build succeeds."""
    from oracle import pyoracle as O
    build = "x86-64-v3 (portable build)"
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode == 0 and os.path.exists(O.NATIVE_LIB_PATH):
        O.use_library(O.NATIVE_LIB_PATH)
        build = "-O3 -march=native -fopenmp, built on this host"
    else:
        log("bench: native oracle build failed, using the portable one:\n" + r.stdout[-2000:])
    threads = O.num_threads()
    rmax = max(d for d in range(1, threads + 1) if n % d == 0)
    dims = (n, n, n)
    V0 = O.solenoidal(dims, SEED, K0, 1.0)
    ns = O.NS(dims, NU, DT)
    ns.set_initial(V0)
    del V0
    t0 = time.perf_counter()
    ns.startup()
    t_start = time.perf_counter() - t0
    sweep = []
    try:
        for ranks in sorted({1, rmax}):
            O.set_ranks(ranks)
            t0 = time.perf_counter()
            ns.steps(steps)
            dt = time.perf_counter() - t0
            sweep.append({"ranks": ranks, "threads_per_rank": threads if ranks == 1 else max(1, threads // ranks),
                          "steps": steps, "steps_per_s": round(steps / dt, 4)})
            log("bench: CPU %d rank(s): %.4f steps/s" % (ranks, steps / dt))
    finally:
        O.set_ranks(1)
        ns.close()
    best = max(sweep, key=lambda s: s["steps_per_s"])
    return best, sweep, threads, t_start, build


TUNE = True


def tune_report(L):
    buf = ctypes.create_string_buffer(8192)
    L.sh_tune_report(buf, 8192)
    return [l for l in buf.value.decode().splitlines() if l]


def main():
    global TUNE
    a = parse()
    TUNE = not a.no_tune
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the multi-GPU slab all-to-all (the reference's rput "
              "transposes, src/Scalar.cpp:486-503, 582-596) is not part of this library; run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import spectral_upcxx_amd as S
    torch.cuda.set_device(0)
    L = S.lib()
    L.sh_init(0)
    n = a.n
    log("bench: %d^3 FP64 fused step, warmup %d, steps %d" % (n, a.warmup, a.steps))
    dt, kern, st = run_ns(S, n, S.SH_F64, a.warmup, a.steps, profile=True)
    value = a.steps / dt
    log("bench: %.2f steps/s (%.3f ms/step), final energy %r, div avg %.2e"
        % (value, 1e3 * dt / a.steps, st["energy"], st["div_avg"]))
    kernels, roof, step_bytes, step_passes = kernel_table(kern, n, st["half"], st["form"])
    state = {"final_energy": st["energy"], "final_div_avg": st["div_avg"],
             "finite": bool(st["energy"] == st["energy"] and abs(st["energy"]) < 1e30),
             "steps_done": a.warmup + a.steps,
             "note": "the reference scheme (explicit AB2 advection, no dealiasing) is unstable at "
                     "256^3, dt=2e-3: the CPU oracle and the GPU overflow between steps 60 and 70; "
                     "per-step cost does not depend on the values"}
    full = None
    if not a.no_full:
        # the plain method: the same step on the full complex spectrum (C2C)
        dtf, kernf, stf = run_ns(S, n, S.SH_F64, a.warmup, a.steps, profile=True, layout=0)
        kf, rooff, bytesf, passesf = kernel_table(kernf, n, stf["half"], stf["form"])
        vf = a.steps / dtf
        full = {"steps_per_s": round(vf, 3), "ms_per_step": round(1e3 * dtf / a.steps, 4),
                "layout": "full spectrum (C2C)", "roofline": rooff, "kernels": kf,
                "step_algo_GB_fused": round(bytesf / 1e9, 3), "step_GBps_fused": round(bytesf * vf / 1e9, 1),
                "step_contract_GB": round(1056.0 * n ** 3 / 1e9, 3),
                "fused_bytes_formula": "%d N^3 B/step (%d field passes)" % (16 * passesf, passesf),
                "history_form": stf["form"],
                "contract_bytes_formula": "1056 N^3 B/step (9 3D FFTs + 12 array passes)",
                "final_energy": stf["energy"]}
        log("bench: full-spectrum (plain C2C) %.2f steps/s" % vf)
    if a.stable_dt > 0:
        dts, _, sts = run_ns(S, n, S.SH_F64, a.warmup, a.stable_steps, profile=False, dt=a.stable_dt)
        state["stable_dt"] = {"dt": a.stable_dt, "steps": a.stable_steps,
                              "steps_per_s": round(a.stable_steps / dts, 3),
                              "final_energy": sts["energy"], "final_div_avg": sts["div_avg"]}
        log("bench: dt=%g: %.2f steps/s, final energy %r" % (a.stable_dt, a.stable_steps / dts, sts["energy"]))
    step = {"chosen_form_algo_GB": round(step_bytes / 1e9, 3),
            "chosen_form_GBps": round(step_bytes * value / 1e9, 1),
            "chosen_form_formula": ("%d N^3 B/step (%d field passes over the N^3/2 stored points)"
                                    % (8 * step_passes, step_passes)) if st["half"] else
                                   ("%d N^3 B/step (%d field passes)" % (16 * step_passes, step_passes)),
            "history_form": st["form"]}
    copy = None
    if not a.no_copy:
        log("bench: HBM copy ceiling")
        copy = copy_ceiling(S)
        log("bench: copy %.0f GB/s" % copy["GBps"])
        if roof:
            roof["frac_of_copy_ceiling"] = round(roof["achieved"] / copy["GBps"], 4)
    fp32 = None
    if a.fp32_steps > 0:
        log("bench: FP32 form, %d steps" % a.fp32_steps)
        dt32, _, _ = run_ns(S, n, S.SH_F32, a.warmup, a.fp32_steps, profile=False, dt=a.stable_dt or DT)
        fp32 = round(a.fp32_steps / dt32, 2)
        log("bench: FP32 %.2f steps/s" % fp32)
    les = None
    if a.les_steps > 0:
        log("bench: 64^3 LES step rate (closure on / off), %d steps per graph" % a.les_steps)
        from spectral_upcxx_amd import forms
        r = forms.les_speed(n=64, steps=a.les_steps, reps=5)
        les = {"grid": 64, "model_steps_per_s": round(r["model_steps_per_s"], 1),
               "model_lagged_steps_per_s": round(r["model_lagged_steps_per_s"], 1),
               "no_model_steps_per_s": round(r["no_model_steps_per_s"], 1),
               "model_vs_chosen_form_dns": round(r["model_steps_per_s"] / value, 2),
               "model_lagged_vs_chosen_form_dns": round(r["model_lagged_steps_per_s"] / value, 2)}
        if full:
            les["model_vs_plain_c2c_dns"] = round(r["model_steps_per_s"] / full["steps_per_s"], 2)
            les["model_lagged_vs_plain_c2c_dns"] = round(r["model_lagged_steps_per_s"] / full["steps_per_s"], 2)
        log("bench: LES %.0f steps/s with the closure (%.1fx the chosen-form DNS), %.0f lagged"
            % (les["model_steps_per_s"], les["model_vs_chosen_form_dns"], les["model_lagged_steps_per_s"]))
    ffts = {}
    if not a.no_fft:
        for nf, reps in ((256, 10), (512, 4)):
            log("bench: 3D FFT %d^3" % nf)
            t, gb = fft_bench(S, nf, reps)
            ffts["fft3d_%d" % nf] = {"ms": round(t * 1e3, 4), "GBps": round(gb, 1),
                                     "frac_hbm": round(gb / HBM_PEAK_GBPS, 4),
                                     "bytes_formula": "96 N^3 B (3 axis passes, read + write, c128)"}
            log("bench: fft %d^3 %.3f ms %.0f GB/s" % (nf, t * 1e3, gb))
    cpu = None
    if not a.no_cpu:
        log("bench: CPU baseline (oracle), start-up + %d steps per rank setting at %d^3" % (a.cpu_steps, n))
        best, sweep, cores, t_start, build = cpu_baseline(n, a.cpu_steps)
        cpu = {"value": best["steps_per_s"], "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": "%d^3 FP64 fractional step (config 2 IC), the oracle's op-by-op restatement of "
                         "exec/ns.cpp:132-217 (project Stockham CPU FFT; FFTW absent): the real start-up "
                         "(%.1f s, untimed), then %d timed steps at each slab-rank setting; best of the "
                         "sweep" % (n, t_start, a.cpu_steps),
               "sweep": sweep, "best_ranks": best["ranks"], "build": build,
               "startup_s": round(t_start, 3),
               "gpu_vs_cpu": round(value / best["steps_per_s"], 1)}
        if full:
            cpu["plain_c2c_gpu_vs_cpu"] = round(full["steps_per_s"] / best["steps_per_s"], 1)
    out = {
        "metric": "timesteps/s at 256^3 FP64 plain vs chosen form; 3D-FFT achieved HBM GB/s",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * dt / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ns_fractional_step_fp64_%d^3" % n, "grid": n, "nu": NU, "dt": DT,
                   "ic": "random solenoidal, E(k)~k^4 exp(-2(k/4)^2), u_rms=1, seed %d" % SEED,
                   "form": ("chosen form: half-spectrum storage of the real fields (planes kz < nz/2, "
                            "R2C-equivalent)%s, %d N^3 B/step; the plain C2C method is full_spectrum"
                            % ({0: "", 1: ", merged AB2 history in kernel E",
                                2: ", merged AB2 history in kernel E, five emitted fields"}[st["form"]], 8 * step_passes))
                   if st["half"] else "plain method: full complex spectrum (C2C), %d N^3 B/step" % (16 * step_passes),
                   "schedule": "fused (4 kernels/step)",
                   "layout": "half spectrum (kz < nz/2; real fields)" if st["half"] else "full spectrum"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "full_spectrum": full,
        "hbm_copy_ceiling": copy,
        "step_bytes": step,
        "chosen_form_fp32_steps_per_s": fp32,
        "les_64": les,
        "fft": ffts,
        "kernels": kernels,
        "state": state,
        "schedules": tune_report(L) if TUNE else "defaults",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

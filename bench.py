"""Benchmark: timesteps/s of the 256^3 FP64 fractional step on one MI355X.

Workload (the headline configuration): plain method, FP64, 256^3 on [0, 2 pi)^3,
random solenoidal IC (E(k) ~ k^4 exp(-2 (k/4)^2), u_rms = 1, seed 2024),
nu = 1/800, dt = 2e-3.  One "step" is one iteration of exec/ns.cpp:167-217
(fused kernel schedule; the start is real, so the solver steps the
half-spectrum layout, the planes kz < nz/2 -- the same velocity to rounding).
Inputs are resident in HBM when the timed region starts.  The same step on
the full complex spectrum (the reference's C2C arithmetic) is reported beside.

Prints ONE JSON line with the metric, a roofline object for the dominant
kernel (algorithmic bytes / average launch time from HIP events recorded on
the launch stream during the timed region), the CPU baseline (the oracle's
op-by-op restatement of the same step on the host cores, bounded sample) and
extra fields: the FP32 form's steps/s and the standalone 3D FFT's HBM GB/s at
256^3 and 512^3 (exec/fft.cpp protocol: min over repeats of forward+inverse).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
NU, DT, SEED, K0 = 1.0 / 800.0, 2e-3, 2024, 4.0
# algorithmic field passes (one complex element read or written = one pass
# unit of 16 B in FP64 / 8 B in FP32, per grid point) of each kernel class
# HBM bytes per launch at 256^3 measured with rocprofv3 PMC (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, factors from a 1 GiB device copy); see
# profiles/pmc_traffic_ns256.md.  Measured for the kernel variants in this tree.
# Keyed by the solver's storage: "full" spectrum or the "half" (kz < nz/2) layout.
MEASURED_TRAFFIC_256 = {
    "full": {"ns_cross_z": 2.416e9, "ns_fourier_x": 5.638e9, "ns_inv_x": 2.416e9,
             "ns_y_inv": 3.221e9, "ns_y_fwd": 1.611e9},
    "half": {"ns_cross_z": 1.208e9, "ns_fourier_x": 2.820e9, "ns_inv_x": 1.208e9,
             "ns_y_inv": 1.611e9, "ns_y_fwd": 0.805e9},
}
PASSES = {
    "ns_inv_x": 3 + 6,        # read V, write 6 x-transformed fields
    "ns_y_inv": 6 + 6,        # y pass of 6 fields
    "ns_cross_z": 6 + 3,      # z inverse of 6, cross, z forward of 3
    "ns_y_fwd": 3 + 3,        # y pass of 3 fields
    "ns_fourier_x": 9 + 12,   # read X, V1, N0; write V1, N0 and next step's 6 x-inverse fields
    "fft_x": 2, "fft_y": 2, "fft_z": 2,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--fp32-steps", type=int, default=100)
    p.add_argument("--les-steps", type=int, default=200, help="0 skips the LES step-rate line")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fft", action="store_true")
    p.add_argument("--no-full", action="store_true", help="skip the full-spectrum FP64 line")
    p.add_argument("--cpu-steps", type=int, default=2)
    p.add_argument("--stable-dt", type=float, default=5e-4)
    p.add_argument("--no-tune", action="store_true", help="use the default schedules")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first run three 1 GiB device-to-device copies (the reference kernel for "
                        "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
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
    ns.steps(warmup)
    torch.cuda.synchronize()
    if profile:
        L.sh_prof_reset()
        L.sh_prof_enable(1)
    t0 = time.perf_counter()
    ns.steps(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern = {}
    if profile:
        L.sh_prof_enable(0)
        for name in ("ns_inv_x", "ns_y_inv", "ns_cross_z", "ns_y_fwd", "ns_fourier_x"):
            ms, cnt = prof_total(L, name)
            kern[name] = (ms, cnt)
    div = ns.div_average()
    V = ns.velocity()
    e = ctypes.c_double()
    L.sh_spectrum_z(S.ptr(V), n, n, n, NU, None, 0, ctypes.byref(e), None, S.stream())
    del V
    ns.close()
    torch.cuda.empty_cache()
    return t1 - t0, kern, {"div_avg": div, "energy": e.value, "half": bool(half.value)}


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


def cpu_baseline(n, steps):
    """The oracle's op-by-op step (exec/ns.cpp:167-217 restated, OpenMP) on the host."""
    from oracle import pyoracle as O
    dims = (n, n, n)
    V0 = O.solenoidal(dims, SEED, K0, 1.0)
    ns = O.NS(dims, NU, DT)
    ns.set_initial(V0)
    del V0
    ns.prime_for_timing()
    t0 = time.perf_counter()
    ns.steps(steps)
    t1 = time.perf_counter()
    ns.close()
    return steps / (t1 - t0), O.num_threads()


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
        print("bench.py runs on one GPU only: the multi-GPU slab all-to-all is not part of this "
              "build (see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import spectral_upcxx_amd as S
    torch.cuda.set_device(0)
    S.lib().sh_init(0)
    n = a.n
    if a.calibrate_copy:
        nb = 1 << 30
        src = torch.empty(nb, dtype=torch.uint8, device="cuda")
        dst = torch.empty(nb, dtype=torch.uint8, device="cuda")
        for _ in range(3):
            S.lib().sh_memcpy_d2d(S.ptr(dst), S.ptr(src), nb, S.stream())
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
        log("bench: calibration copies done (3 x 1 GiB)")
    log("bench: %d^3 FP64 fused step, warmup %d, steps %d" % (n, a.warmup, a.steps))
    dt, kern, st = run_ns(S, n, S.SH_F64, a.warmup, a.steps, profile=True)
    value = a.steps / dt
    log("bench: %.2f steps/s (%.3f ms/step), final energy %r, div avg %.2e"
        % (value, 1e3 * dt / a.steps, st["energy"], st["div_avg"]))
    state = {"final_energy": st["energy"], "final_div_avg": st["div_avg"],
             "finite": bool(st["energy"] == st["energy"] and abs(st["energy"]) < 1e30),
             "note": "the reference scheme (explicit AB2 advection, no dealiasing) is unstable at "
                     "256^3, dt=2e-3: the CPU oracle also overflows near step 65; per-step cost "
                     "does not depend on the values"}
    full = None
    if not a.no_full:
        # the same step on the full complex spectrum (the reference's C2C arithmetic)
        dtf, _, _ = run_ns(S, n, S.SH_F64, a.warmup, a.steps, profile=False, layout=0)
        full = round(a.steps / dtf, 3)
        log("bench: full-spectrum layout %.2f steps/s" % full)
    if a.stable_dt > 0:
        dts, _, sts = run_ns(S, n, S.SH_F64, a.warmup, a.steps, profile=False, dt=a.stable_dt)
        state["stable_dt"] = {"dt": a.stable_dt, "steps_per_s": round(a.steps / dts, 3),
                              "final_energy": sts["energy"], "final_div_avg": sts["div_avg"]}
        log("bench: dt=%g: %.2f steps/s, final energy %r" % (a.stable_dt, a.steps / dts, sts["energy"]))
    # grid points each kernel touches per field: the half-spectrum layout
    # stores the planes kz < nz/2 only
    layout = "half" if st["half"] else "full"
    pts = float(n) ** 3 * (0.5 if st["half"] else 1.0)
    kernels = {}
    dom, dom_t = None, -1.0
    for name, (ms, cnt) in kern.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        byts = PASSES[name] * 16.0 * pts
        kernels[name] = {"avg_ms": round(avg_ms, 5), "launches": cnt,
                         "algo_GB": round(byts / 1e9, 4), "GBps": round(byts / (avg_ms * 1e-3) / 1e9, 1)}
        if ms > dom_t:
            dom, dom_t = name, ms
    roof = None
    if dom:
        k = kernels[dom]
        ach = k["GBps"]
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBPS, 4),
                "traffic": MEASURED_TRAFFIC_256[layout].get(dom) if n == 256 else None,
                "algo_bytes_per_launch": PASSES[dom] * 16.0 * pts}
    step_bytes = sum(PASSES[k] for k, (ms, cnt) in kern.items() if cnt > 0) * 16.0 * pts
    extra = {"step_algo_GB_fused": round(step_bytes / 1e9, 3),
             "step_GBps_fused": round(step_bytes * value / 1e9, 1),
             "step_contract_GB": round(1056.0 * n ** 3 / 1e9, 3),
             "step_GBps_contract": round(1056.0 * n ** 3 * value / 1e9, 1)}
    fp32 = None
    if a.fp32_steps > 0:
        log("bench: FP32 form, %d steps" % a.fp32_steps)
        dt32, _, _ = run_ns(S, n, S.SH_F32, a.warmup, a.fp32_steps, profile=False)
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
               "model_vs_fp64_dns_steps": round(r["model_steps_per_s"] / value, 2),
               "model_lagged_vs_fp64_dns_steps": round(r["model_lagged_steps_per_s"] / value, 2)}
        log("bench: LES %.0f steps/s with the closure (%.1fx the FP64 step rate), %.0f with the lagged "
            "closure (%.1fx)" % (les["model_steps_per_s"], les["model_vs_fp64_dns_steps"],
                                 les["model_lagged_steps_per_s"], les["model_lagged_vs_fp64_dns_steps"]))
    ffts = {}
    if not a.no_fft:
        for nf, reps in ((256, 10), (512, 4)):
            log("bench: 3D FFT %d^3" % nf)
            t, gb = fft_bench(S, nf, reps)
            ffts["fft3d_%d" % nf] = {"ms": round(t * 1e3, 4), "GBps": round(gb, 1),
                                     "frac_hbm": round(gb / HBM_PEAK_GBPS, 4)}
            log("bench: fft %d^3 %.3f ms %.0f GB/s" % (nf, t * 1e3, gb))
    cpu = None
    if not a.no_cpu:
        log("bench: CPU baseline (oracle) %d steps at %d^3" % (a.cpu_steps, n))
        v, cores = cpu_baseline(n, a.cpu_steps)
        cpu = {"value": round(v, 4), "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": "%d steps of the %d^3 FP64 fractional step, oracle op-by-op restatement of "
                         "exec/ns.cpp:167-217 (project Stockham FFT, FFTW absent), OpenMP %d threads"
                         % (a.cpu_steps, n, cores)}
        log("bench: CPU %.4f steps/s on %d threads" % (v, cores))
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
                   "schedule": "fused (4 kernels/step)",
                   "layout": "half spectrum (kz < nz/2; real fields)" if st["half"] else "full spectrum"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "full_spectrum_fp64_steps_per_s": full,
        "chosen_form_fp32_steps_per_s": fp32,
        "les_64": les,
        "fft": ffts,
        "kernels": kernels,
        "bytes": extra,
        "state": state,
        "schedules": tune_report(S.lib()) if TUNE else "defaults",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Harnesses for the solver's forms, driving the C ABI (device buffers via torch).

  python -m spectral_upcxx_amd.forms fp32 --n 512 --steps 100 --dt 2.5e-4

fp32: the FP32 solver (complex64 transforms and operators; FP64 twiddles, CN
factors and reductions) against the FP64 plain method from the same seed.
Reports the relative L2 velocity error, the shell-spectrum error and the
dissipation-rate error, and both solvers' steps/s.
"""
import argparse
import ctypes
import json
import sys
import time

from . import lib, ptr, stream, NS, SH_F32, SH_F64, SH_NORMALIZE


def _torch():
    import torch
    return torch


def solenoidal(n, seed, k0=4.0, urms=1.0):
    torch = _torch()
    V = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
    lib().sh_solenoidal_ic_z(ptr(V), n, n, n, int(seed), float(k0), float(urms), stream())
    return V


def spectrum(V, nu, nbins):
    """Shell spectrum E(k), total energy and dissipation 2 nu sum k^2 E (FP64)."""
    nz, ny, nx = V.shape[-3:]
    Ek = (ctypes.c_double * nbins)()
    e = ctypes.c_double()
    d = ctypes.c_double()
    lib().sh_spectrum_z(ptr(V), nx, ny, nz, float(nu), Ek, nbins, ctypes.byref(e), ctypes.byref(d), stream())
    return list(Ek), e.value, d.value


def run_solver(V0, nu, dt, steps, precision, tune=True):
    torch = _torch()
    n = V0.shape[-1]
    ns = NS((n, n, n), nu, dt, precision, fused=True)
    ns.set_initial(V0)
    ns.startup()
    if tune:
        lib().sh_ns_autotune(ns.h, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ns.steps(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    V = ns.velocity()
    ns.close()
    lib().sh_pool_trim()
    return V, steps / (t1 - t0)


def fp32_error(n=512, steps=100, dt=2.5e-4, nu=1.0 / 800, seed=2024, k0=4.0):
    torch = _torch()
    V0 = solenoidal(n, seed, k0)
    V64, sps64 = run_solver(V0, nu, dt, steps, SH_F64)
    V32, sps32 = run_solver(V0, nu, dt, steps, SH_F32)
    del V0
    nb = n // 2
    E64, e64, d64 = spectrum(V64, nu, nb)
    E32, e32, d32 = spectrum(V32, nu, nb)
    rel_v = float(torch.linalg.vector_norm(V32 - V64) / torch.linalg.vector_norm(V64))
    shells = [k for k in range(1, nb) if E64[k] > 1e-12 * max(E64)]
    rel_ek = [abs(E32[k] - E64[k]) / E64[k] for k in shells]
    k16 = [abs(E32[k] - E64[k]) / E64[k] for k in shells if k <= 16]
    del V64, V32
    torch.cuda.empty_cache()
    return {
        "form": "fp32", "grid": n, "steps": steps, "dt": dt, "nu": nu, "seed": seed, "k0": k0,
        "velocity_rel_l2": rel_v,
        "spectrum_max_rel_err": max(rel_ek), "spectrum_max_rel_err_k_le_16": max(k16),
        "spectrum_shells_compared": len(shells),
        "energy_fp64": e64, "energy_fp32": e32, "energy_rel_err": abs(e32 - e64) / e64,
        "dissipation_fp64": d64, "dissipation_fp32": d32, "dissipation_rel_err": abs(d32 - d64) / d64,
        "steps_per_s_fp64": sps64, "steps_per_s_fp32": sps32,
    }


def _u01(seed, counter):
    """splitmix64 draw (the PRNG of the ICs), U[0,1)."""
    m = (1 << 64) - 1
    z = (seed + (counter + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) * 2.0 ** -53


def ensemble_params(groups, seed):
    """Per-box (seed, k0 ~ U[2,6], nu ~ U[1/1000, 1/300]) for groups [(n, nboxes)]."""
    out = []
    c = 0
    for n, nb in groups:
        boxes = []
        for b in range(nb):
            k0 = 2.0 + 4.0 * _u01(seed, 2 * c)
            nu = 1.0 / 1000 + (1.0 / 300 - 1.0 / 1000) * _u01(seed, 2 * c + 1)
            boxes.append({"seed": seed * 1000 + c, "k0": k0, "nu": nu})
            c += 1
        out.append((n, boxes))
    return out


class Ensemble:
    """Groups of equal boxes (one sh_ns ensemble handle per size) on one stream."""

    def __init__(self, groups, dt, precision=SH_F32, seed=4242, tune=True):
        torch = _torch()
        self.L = lib()
        self.stream = ctypes.c_void_p()
        self.L.sh_stream_create(ctypes.byref(self.stream))
        self.params = ensemble_params(groups, seed)
        self.handles = []
        self.dt = dt
        self.precision = precision
        for n, boxes in self.params:
            h = ctypes.c_void_p()
            nus = (ctypes.c_double * len(boxes))(*[b["nu"] for b in boxes])
            self.L.sh_ns_create_ensemble(ctypes.byref(h), n, len(boxes), nus, float(dt), int(precision), self.stream)
            V0 = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
            for i, b in enumerate(boxes):
                self.L.sh_solenoidal_ic_z(ptr(V0), n, n, n, b["seed"], b["k0"], 1.0, self.stream)
                self.L.sh_ns_set_initial_box(h, i, ptr(V0))
            self.L.sh_stream_synchronize(self.stream)
            del V0
            if tune:
                self.L.sh_ns_autotune(h, 2)
            self.handles.append(h)

    def steps(self, k):
        for h in self.handles:
            self.L.sh_ns_steps(h, int(k))

    def snapshot_shapes(self, ratio):
        return [(len(boxes), 9, (n // ratio) ** 3) for n, boxes in self.params]

    def snapshot(self, bufs, ratio):
        for h, b in zip(self.handles, bufs):
            self.L.sh_ns_snapshot(h, int(ratio), ptr(b))

    def capture(self, nsteps, every, ratio, snaps):
        """One graph: nsteps steps of every group with a snapshot after each `every`.
        Each handle's state is first put in the form its tuned schedule steps
        (sh_ns_prepare), so the graph records the merged-history kernels."""
        for h in self.handles:
            if self.L.sh_ns_prepare(h) != 0:
                raise RuntimeError(self.L.sh_last_error().decode())
        self.L.sh_graph_begin(self.stream)
        try:
            for k in range(nsteps // every):
                self.steps(every)
                self.snapshot([s[k] for s in snaps], ratio)
        except Exception:
            g = ctypes.c_void_p()
            self.L.sh_graph_end(self.stream, ctypes.byref(g))
            raise
        g = ctypes.c_void_p()
        self.L.sh_graph_end(self.stream, ctypes.byref(g))
        return g

    def velocity(self, group, box):
        torch = _torch()
        n = self.params[group][0]
        V = torch.empty((3, n, n, n), dtype=torch.complex128, device="cuda")
        self.L.sh_ns_get_velocity_box(self.handles[group], box, ptr(V))
        self.L.sh_stream_synchronize(self.stream)
        return V

    def close(self):
        for h in self.handles:
            self.L.sh_ns_destroy(h)
        self.handles = []
        if self.stream:
            self.L.sh_stream_destroy(self.stream)
            self.stream = ctypes.c_void_p()


def write_snapshots(path, ens, snaps, every, ratio):
    """Training-data file: a JSON header line, then float32 arrays in the
    order of the header's `arrays` list (C order)."""
    import numpy as np
    header = {"format": "spectral-snapshots-v1", "ratio": ratio, "every": every, "dt": ens.dt,
              "fields": ["ubar_x", "ubar_y", "ubar_z", "tau_xx", "tau_yy", "tau_zz", "tau_xy", "tau_xz", "tau_yz"],
              "arrays": []}
    for (n, boxes), s in zip(ens.params, snaps):
        header["arrays"].append({"n": n, "coarse": n // ratio, "boxes": boxes,
                                 "shape": [int(s.shape[0]), len(boxes), 9] + [n // ratio] * 3})
    with open(path, "wb") as f:
        f.write((json.dumps(header) + "\n").encode())
        for s in snaps:
            f.write(np.ascontiguousarray(s.cpu().numpy(), dtype=np.float32).tobytes())


def read_snapshots(path):
    import numpy as np
    with open(path, "rb") as f:
        header = json.loads(f.readline().decode())
        arrays = []
        for a in header["arrays"]:
            cnt = int(np.prod(a["shape"]))
            arrays.append(np.frombuffer(f.read(4 * cnt), dtype=np.float32).reshape(a["shape"]))
    return header, arrays


def ensemble_run(groups=((64, 32), (96, 16)), steps=500, every=10, ratio=4, dt=1e-3,
                 precision=SH_F32, seed=4242, out=None):
    torch = _torch()
    ens = Ensemble(groups, dt, precision, seed)
    nsnap = steps // every
    snaps = [torch.empty((nsnap,) + shp, dtype=torch.float32, device="cuda") for shp in ens.snapshot_shapes(ratio)]
    ens.snapshot([s[0] for s in snaps], ratio)  # allocates the snapshot scratch before capture
    t0 = time.perf_counter()
    g = ens.capture(steps, every, ratio, snaps)
    t1 = time.perf_counter()
    lib().sh_graph_launch(g, ens.stream)
    lib().sh_stream_synchronize(ens.stream)
    t2 = time.perf_counter()
    lib().sh_graph_destroy(g)
    boxes = sum(nb for _, nb in groups)
    energies = []
    for gi, (n, bx) in enumerate(ens.params):
        for b in (0, len(bx) - 1):
            V = ens.velocity(gi, b)
            _, e, _ = spectrum(V, bx[b]["nu"], 2)
            energies.append({"n": n, "box": b, "energy": e})
    res = {"form": "ensemble", "groups": [list(x) for x in groups], "boxes": boxes, "steps": steps,
           "snapshot_every": every, "ratio": ratio, "dt": dt, "precision": precision,
           "capture_s": t1 - t0, "graph_run_s": t2 - t1,
           "box_steps_per_s": boxes * steps / (t2 - t1),
           "grid_points_steps_per_s": sum(n ** 3 * nb for n, nb in groups) * steps / (t2 - t1),
           "sample_energies": energies}
    if out:
        write_snapshots(out, ens, snaps, every, ratio)
        res["snapshots_file"] = out
    ens.close()
    del snaps
    torch.cuda.empty_cache()
    return res


def physical_to_fourier(u):
    """[3, n, n, n] float32 physical field -> complex128 Fourier fields (the
    solver's unnormalised forward transform)."""
    torch = _torch()
    n = u.shape[-1]
    V = u.to(torch.complex128).contiguous()
    lib().sh_fft3d_many_z2z(ptr(V), 3, n ** 3, n, n, n, -1, 0, stream())
    return V


def project_fourier(V):
    """In place: zero the Nyquist modes of Fourier fields V (complex128
    [3, n, n, n]) and remove their gradient part (V - grad lap^-1 div V)."""
    torch = _torch()
    L = lib()
    n = V.shape[-1]
    for i in range(3):
        L.sh_zero_high_z(ptr(V[i]), n, n, n, stream())
    d = torch.empty((n, n, n), dtype=torch.complex128, device=V.device)
    L.sh_divergence_z(ptr(V[0]), ptr(V[1]), ptr(V[2]), ptr(d), n, n, n, stream())
    L.sh_laplacian_inverse_z(ptr(d), 0.0, 1.0, n, n, n, stream())
    G = torch.empty_like(V)
    L.sh_gradient_z(ptr(d), ptr(G[0]), ptr(G[1]), ptr(G[2]), n, n, n, stream())
    L.sh_sub_z(ptr(V), ptr(V), ptr(G), 3 * n ** 3, stream())
    return V


def closure_target(ubar, A):
    """Training pair of the force-matching loss from one filtered DNS
    snapshot: the network inputs X ([9, n^3] float32) of the projected
    filtered velocity and the force the coarse solver lacks,
    ft = F(box(u x omega)) - N(ubar) (complex64 [3, n, n, n]), N being the
    coarse solver's own nonlinear term of the projected ubar."""
    torch = _torch()
    from . import closure
    L = lib()
    n = ubar.shape[-1]
    V = project_fourier(physical_to_fourier(ubar))
    up = V.clone()
    L.sh_fft3d_many_z2z(ptr(up), 3, n ** 3, n, n, n, 1, SH_NORMALIZE, stream())
    X = closure.features(up.real.float().contiguous())
    W = torch.empty_like(V)
    L.sh_curl_z(ptr(V[0]), ptr(V[1]), ptr(V[2]), ptr(W[0]), ptr(W[1]), ptr(W[2]), n, n, n, stream())
    Nl = torch.empty_like(V)
    work = torch.empty((6, n, n, n), dtype=torch.complex128, device=V.device)
    L.sh_cross_z(ptr(V), ptr(W), ptr(Nl), ptr(work), n, n, n, stream())
    Ah = physical_to_fourier(A)
    L.sh_sub_z(ptr(Ah), ptr(Ah), ptr(Nl), 3 * n ** 3, stream())
    return X, Ah.to(torch.complex64).contiguous()


def _spectrum_rel(Ea, Eb, kmax):
    return [abs(Ea[k] - Eb[k]) / Eb[k] for k in range(1, kmax + 1)]


def les_study(n_dns=256, ratio=4, nu=1.0 / 800, dt_dns=5e-4, les_dt=2e-3, t_end=2.0, snap_dt=0.25,
              train_seeds=8, test_seeds=4, seed0=7100, k0=4.0, iters=3000, batch=65536, lr=2e-3,
              variants=((True, 1), (False, 1), (True, 2), (False, 2)), eval_dt=0.5, kcut=16, net_seed=5,
              gain_candidates=(0.5, 0.75, 1.0, 1.25, 1.5, 2.0, 3.0), force_iters=2000, force_lr=1e-3,
              force_variants=(("force", False, 0.0, 1), ("force_exp", False, 0.0, 2)),
              cs_candidates=(0.0, 0.05, 0.1), select="mean", log=None):
    """Learned-closure form: FP64 DNS at n_dns^3 box-filtered by `ratio`
    gives training pairs (filtered velocity gradients -> stresses); the
    network is trained on the GPU; the FP32 LES on the coarse grid with and
    without the model is compared with the filtered DNS spectrum for k <=
    kcut on held-out seeds.  The network variant, clip, dealiasing, gain and
    C_s are chosen on two training seeds by the mean over the evaluation times
    of the mean (select="mean") or the maximum (select="max") over k <= kcut
    of the relative E(k) error."""
    import math
    torch = _torch()
    from . import closure
    log = log or (lambda *a: print(*a, file=sys.stderr, flush=True))
    L = lib()
    n = n_dns // ratio
    nc = n ** 3
    every = int(round(les_dt / dt_dns))
    assert abs(every * dt_dns - les_dt) < 1e-12, "the LES step must be a multiple of the DNS step"
    snap_steps = int(round(snap_dt / dt_dns))
    eval_steps = int(round(eval_dt / dt_dns))
    total = int(round(t_end / dt_dns))
    assert snap_steps % every == 0 and eval_steps % every == 0 and total % eval_steps == 0
    # ---- DNS: one solver reused for every seed
    dns = NS((n_dns,) * 3, nu, dt_dns, SH_F64, fused=True)
    snap = torch.empty((1, 9, nc), dtype=torch.float32, device="cuda")
    fsnap = torch.empty((1, 3, nc), dtype=torch.float32, device="cuda")
    train, test = [], []
    dns_time = 0.0
    dns_steps = 0
    tuned = False
    for si in range(train_seeds + test_seeds):
        seed = seed0 + si
        is_train = si < train_seeds
        V0 = solenoidal(n_dns, seed, k0)
        dns.set_initial(V0)
        del V0
        dns.startup()
        if not tuned:
            L.sh_ns_autotune(dns.h, 2)
            tuned = True
        L.sh_ns_snapshot(dns.h, ratio, ptr(snap))
        snaps = {0: snap[0].clone()}
        forces = {}
        chunk = snap_steps if is_train else eval_steps
        done = 0
        while done < total:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dns.steps(chunk)
            torch.cuda.synchronize()
            dns_time += time.perf_counter() - t0
            dns_steps += chunk
            done += chunk
            L.sh_ns_snapshot(dns.h, ratio, ptr(snap))
            snaps[done] = snap[0].clone()
            L.sh_ns_snapshot_force(dns.h, ratio, ptr(fsnap))
            forces[done] = fsnap[0].clone()
        finite = all(bool(torch.isfinite(s).all()) for s in snaps.values())
        (train if is_train else test).append({"seed": seed, "snaps": snaps, "forces": forces, "finite": finite})
        log("dns seed %d (%s): %d steps, finite=%s, %.1f steps/s" % (
            seed, "train" if is_train else "test", total, finite, dns_steps / dns_time))
    dns.close()
    L.sh_pool_trim()
    dns_sps = dns_steps / dns_time
    # ---- training pairs: every training snapshot after t = 0
    keys = [k for k in sorted(train[0]["snaps"]) if k > 0]
    M = len(train) * len(keys) * nc
    X = torch.empty((9, M), dtype=torch.float32, device="cuda")
    T = torch.empty((6, M), dtype=torch.float32, device="cuda")
    off = 0
    for tr in train:
        for k in keys:
            s = tr["snaps"][k]
            ub = s[:3].reshape(3, n, n, n).contiguous()
            L.sh_les_features(n, ptr(ub), closure._offset(X, off), M, stream())
            T[:, off:off + nc] = s[3:]
            off += nc
    in_scale = (1.0 / X.std(dim=1)).cpu().numpy()
    out_scale = T.std(dim=1).cpu().numpy()
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    perm = torch.randperm(M, device="cuda", generator=g)
    X = X[:, perm].contiguous()
    T = T[:, perm].contiguous()
    del perm
    model = closure.MLP(seed=net_seed)
    model.set_scaling(in_scale, out_scale)
    nbatch = M // batch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = []
    for it in range(iters):
        lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * it / iters)))
        want = it % 500 == 0 or it == iters - 1
        l = model.train_step(X, T, lr_t, n=batch, start=(it % nbatch) * batch, want_loss=want)
        if want:
            losses.append((it, l))
            log("train it %d loss %.4f" % (it, l))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del X, T
    torch.cuda.empty_cache()
    # ---- force matching: fine-tune so that -div tau, as the dealiased coarse
    # solver applies it, supplies the force that solver lacks (gradients go
    # back through the spectral divergence into the network's backward kernel)
    force_log = {}
    candidates = [("a_priori", model.params())]
    if force_iters > 0:
        import random

        def pairs(runs, ks):
            out = []
            for r in runs:
                for k in ks:
                    ub = r["snaps"][k][:3].reshape(3, n, n, n).contiguous()
                    out.append(closure_target(ub, r["forces"][k].reshape(3, n, n, n).contiguous()))
            return out

        fit = pairs(train, keys)
        held = pairs(test, [k for k in sorted(test[0]["forces"])])
        # each field's own target loss (that of a network that outputs zero):
        # the relative variant divides by it, and the held-out check reports it
        zero = closure.MLP(seed=1)
        zero.set_params([0.0] * closure.param_count())
        zero.set_scaling(in_scale, out_scale)
        b = ctypes.c_double()
        for pp in (fit, held):
            for i, (Xs, fts) in enumerate(pp):
                L.sh_mlp_force_gradient(zero.h, n, ptr(Xs), ptr(fts), 1, 1.0, 0.0, None, ctypes.byref(b))
                pp[i] = (Xs, fts, 1.0 / b.value)
        zero.close()

        def rel_loss(m, pp):
            a = ctypes.c_double()
            tot = 0.0
            for Xs, fts, w in pp:
                L.sh_mlp_force_gradient(m.h, n, ptr(Xs), ptr(fts), 1, w, 0.0, None, ctypes.byref(a))
                tot += a.value
            return tot / len(pp)

        force_log["held_out_rel_loss_a_priori"] = rel_loss(model, held)
        base = model.params()
        for name, relative, kw, mode in force_variants:
            # a fresh optimiser state from the a-priori weights for each variant
            m = closure.MLP(seed=net_seed)
            m.set_params(base)
            m.set_scaling(in_scale, out_scale)
            order = list(range(len(fit)))
            rng = random.Random(7)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for it in range(force_iters):
                if it % len(order) == 0:
                    rng.shuffle(order)
                Xs, fts, w = fit[order[it % len(order)]]
                lr_t = force_lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * it / force_iters)))
                want = it % 500 == 0 or it == force_iters - 1
                lv = ctypes.c_double()
                L.sh_mlp_force_train_step(m.h, n, ptr(Xs), ptr(fts), mode, w if relative else 1.0, kw, lr_t,
                                          ctypes.byref(lv) if want else None)
                if want:
                    log("force %s it %d loss %.4e" % (name, it, lv.value))
            torch.cuda.synchronize()
            force_log[name] = {"relative": relative, "k_weight": kw, "dealias": mode,
                               "train_s": time.perf_counter() - t0,
                               "held_out_rel_loss": rel_loss(m, held)}
            log("force matching %s: held-out relative loss %.4f (a priori %.4f)" % (
                name, force_log[name]["held_out_rel_loss"], force_log["held_out_rel_loss_a_priori"]))
            candidates.append((name, m.params()))
            m.close()
        del fit, held
    for r in train + test:
        r["forces"] = {}
    torch.cuda.empty_cache()
    # ---- a-priori check on held-out snapshots: correlation of the stresses
    corr = []
    for te in test:
        s = te["snaps"][total]
        ub = s[:3].reshape(3, n, n, n).contiguous()
        Xt = closure.features(ub)
        Y = model.forward(Xt)
        for c in range(6):
            a = Y[c].double() - Y[c].double().mean()
            b = s[3 + c].double() - s[3 + c].double().mean()
            corr.append(float((a * b).sum() / (a.norm() * b.norm())))
    # ---- a-posteriori: LES with and without the model against the filtered DNS
    nb = n // 2
    kcut = min(kcut, nb - 1)
    targets = list(range(eval_steps, total + 1, eval_steps))
    key = "max_rel_err_k_le_%d" % kcut

    def fdns_spectra(snaps):
        out = {}
        for t in targets:
            uf = snaps[t][:3].reshape(3, n, n, n).contiguous()
            out[t] = spectrum(physical_to_fourier(uf), nu, nb)
        return out

    def trajectory(ub0, mdl, clip, dealias, lag=0, cs=0.0):
        les = closure.LES(n, nu, les_dt, mdl, clip=clip, dealias=dealias, lag=lag)
        if mdl is not None and cs > 0:
            les.set_smagorinsky(cs)
        les.set_initial(ub0)
        cur = every  # after the start-up step the LES is at DNS step `every`
        Es = {}
        for t in targets:
            les.steps((t - cur) // every)
            cur = t
            Es[t] = spectrum(les.velocity(), nu, nb)
        les.close()
        return Es

    def compare(Es, Fs):
        rows = []
        for t in targets:
            E, e, d = Es[t]
            Ef, ef, df = Fs[t]
            finite = all(math.isfinite(x) for x in E[:kcut + 1]) and math.isfinite(e)
            rel = _spectrum_rel(E, Ef, kcut) if finite else [float("inf")] * kcut
            rows.append({"t": t * dt_dns, key: max(rel), "mean_rel_err_k_le_%d" % kcut: sum(rel) / kcut,
                         "energy": e, "energy_fdns": ef, "dissipation": d,
                         "ratio_by_k": [round(E[k] / Ef[k], 4) if finite else None for k in range(1, kcut + 1)]})
        return rows

    def score(rows):
        return sum(r["mean_rel_err_k_le_%d" % kcut] for r in rows) / len(rows)

    def select_score(rows):
        return score(rows) if select == "mean" else sum(r[key] for r in rows) / len(rows)

    # schedules for the coarse grid (kept per process)
    les = closure.LES(n, nu, les_dt, model, clip=True)
    les.set_initial(train[0]["snaps"][0][:3].reshape(3, n, n, n).contiguous())
    les.autotune(2)
    les.close()
    # a-posteriori choices -- the network (a-priori or force-matched),
    # backscatter clipping, the coarse solver's dealiasing mode, the gain of
    # the network's output -- made on training
    # seeds only; the held-out seeds are evaluated once with the choice
    calib = []
    for tr in train[:2]:
        calib.append((tr["snaps"][0][:3].reshape(3, n, n, n).contiguous(), fdns_spectra(tr["snaps"])))
    scores = {}
    params = dict(candidates)
    for cname, cparams in candidates:
        model.set_params(cparams)
        for clip_v, dealias_v in variants:
            for gain in gain_candidates:
                model.set_scaling(in_scale, out_scale * gain)
                for cs in cs_candidates:
                    s = sum(select_score(compare(trajectory(ub, model, clip_v, dealias_v, cs=cs), Fs))
                            for ub, Fs in calib) / len(calib)
                    scores[(cname, clip_v, dealias_v, gain, cs)] = s
            best_c = min((k for k in scores if k[:3] == (cname, clip_v, dealias_v)), key=scores.get)
            log("calibration %s clip=%d dealias=%d: best gain %.2f Cs %.2f score %.4f" % (
                cname, clip_v, dealias_v, best_c[3], best_c[4], scores[best_c]))
    for dealias_v in sorted({d for _, d in variants}):
        s = sum(select_score(compare(trajectory(ub, None, False, dealias_v), Fs)) for ub, Fs in calib) / len(calib)
        scores[("no_model", dealias_v)] = s
        log("calibration no-model dealias=%d: score %.4f" % (dealias_v, s))
    best = min((k for k in scores if k[0] != "no_model"), key=scores.get)
    chosen, clip, dealias, best_gain, best_cs = best
    log("chosen: %s clip=%d dealias=%d gain %.2f Cs %.2f" % best)
    # the best choice without the eddy-viscosity term, for comparison
    best0 = min((k for k in scores if k[0] != "no_model" and k[4] == 0), key=scores.get)
    model.set_params(params[chosen])
    model.set_scaling(in_scale, out_scale * best_gain)
    for tr in train:
        tr["snaps"] = {0: tr["snaps"][0]}
    evals = []
    for te in test:
        ub0 = te["snaps"][0][:3].reshape(3, n, n, n).contiguous()
        Fs = fdns_spectra(te["snaps"])
        res = {"seed": te["seed"]}
        if best_cs > 0:
            model.set_params(params[best0[0]])
            model.set_scaling(in_scale, out_scale * best0[3])
            res["model_cs0"] = compare(trajectory(ub0, model, best0[1], best0[2]), Fs)
            model.set_params(params[chosen])
            model.set_scaling(in_scale, out_scale * best_gain)
        res.update({"model": compare(trajectory(ub0, model, clip, dealias, cs=best_cs), Fs),
               "model_lagged": compare(trajectory(ub0, model, clip, dealias, lag=1, cs=best_cs), Fs),
               "no_model": compare(trajectory(ub0, None, False, dealias), Fs),
               "no_model_plain": compare(trajectory(ub0, None, False, False), Fs)})
        evals.append(res)
        log("les seed %d: " % te["seed"] + ", ".join(
            "t=%.2f model %.3f lagged %.3f no-model %.3f" % (a["t"], a[key], c[key], b[key])
            for a, b, c in zip(res["model"], res["no_model"], res["model_lagged"])))
    # ---- speed: LES steps/s with each step's launches replayed from one graph
    speed = {}
    ub0 = test[0]["snaps"][0][:3].reshape(3, n, n, n).contiguous()
    st = ctypes.c_void_p()
    L.sh_stream_create(ctypes.byref(st))
    for name, mdl, lag in (("model", model, 0), ("model_lagged", model, 1), ("no_model", None, 0)):
        m2 = None
        if mdl is not None:
            m2 = closure.MLP(seed=net_seed, strm=st)
            m2.set_params(model.params())
            m2.set_scaling(in_scale, out_scale * best_gain)
        les = closure.LES(n, nu, les_dt, m2, clip=clip, strm=st, dealias=dealias, lag=lag)
        if m2 is not None and best_cs > 0:
            les.set_smagorinsky(best_cs)
        torch.cuda.synchronize()
        les.set_initial(ub0)
        les.steps(2)
        L.sh_stream_synchronize(st)
        nrep, per = 5, 100
        L.sh_graph_begin(st)
        les.steps(per)
        gr = ctypes.c_void_p()
        L.sh_graph_end(st, ctypes.byref(gr))
        L.sh_graph_launch(gr, st)
        L.sh_stream_synchronize(st)
        t0 = time.perf_counter()
        for _ in range(nrep):
            L.sh_graph_launch(gr, st)
        L.sh_stream_synchronize(st)
        speed[name] = nrep * per / (time.perf_counter() - t0)
        L.sh_graph_destroy(gr)
        les.close()
        if m2:
            m2.close()
    L.sh_stream_destroy(st)
    model_params = model.params()
    model.close()
    return {
        "form": "les", "select": select, "dns_grid": n_dns, "les_grid": n, "ratio": ratio, "nu": nu, "dt_dns": dt_dns,
        "dt_les": les_dt, "t_end": t_end, "k0": k0, "train_seeds": [t["seed"] for t in train],
        "test_seeds": [t["seed"] for t in test], "dns_all_finite": all(t["finite"] for t in train + test),
        "training_samples": M, "iters": iters, "batch": batch, "lr": lr, "train_s": train_s,
        "smagorinsky_cs": best_cs, "best_without_smagorinsky": str(best0),
        "loss_history": losses, "force_matching": force_log, "network": chosen, "clip": clip, "dealias": dealias,
        "a_priori_corr_by_component": [round(c, 4) for c in corr[:6]],
        "a_priori_corr_mean": sum(corr) / len(corr),
        "calibration_scores": {str(k): v for k, v in scores.items()}, "best_gain": best_gain,
        "evals": evals,
        "final_%s_model_mean" % key: sum(r["model"][-1][key] for r in evals) / len(evals),
        "final_%s_no_model_mean" % key: sum(r["no_model"][-1][key] for r in evals) / len(evals),
        "final_%s_model_lagged_mean" % key: sum(r["model_lagged"][-1][key] for r in evals) / len(evals),
        "mean_rel_err_model": sum(score(r["model"]) for r in evals) / len(evals),
        "mean_rel_err_model_lagged": sum(score(r["model_lagged"]) for r in evals) / len(evals),
        "mean_rel_err_model_cs0": (sum(score(r["model_cs0"]) for r in evals) / len(evals)
                                   if best_cs > 0 else None),
        "final_%s_model_cs0_mean" % key: (sum(r["model_cs0"][-1][key] for r in evals) / len(evals)
                                          if best_cs > 0 else None),
        "mean_rel_err_no_model": sum(score(r["no_model"]) for r in evals) / len(evals),
        "mean_rel_err_no_model_plain": sum(score(r["no_model_plain"]) for r in evals) / len(evals),
        "dns_fp64_steps_per_s": dns_sps, "les_model_steps_per_s": speed["model"],
        "les_no_model_steps_per_s": speed["no_model"],
        "les_model_lagged_steps_per_s": speed["model_lagged"],
        "les_lagged_vs_dns_steps_ratio": speed["model_lagged"] / dns_sps,
        "les_vs_dns_steps_ratio": speed["model"] / dns_sps,
        "les_vs_dns_sim_time_ratio": speed["model"] * les_dt / (dns_sps * dt_dns),
        "param_checksum": float(abs(model_params).sum()),
    }


def les_speed(n=64, steps=200, reps=5, clip=True, seed=7100, nu=1.0 / 800, dt=2e-3):
    """LES step rate with the synchronous closure, with the lagged closure
    (sh_les_set_lag) and without a closure (untrained network with a small
    output scale; the cost does not depend on the weights), each `steps` steps
    replayed from one hipGraph."""
    torch = _torch()
    from . import closure
    L = lib()
    V = solenoidal(n, seed)
    L.sh_fft3d_many_z2z(ptr(V), 3, n ** 3, n, n, n, 1, SH_NORMALIZE, stream())
    ub = V.real.float().contiguous()
    del V
    st = ctypes.c_void_p()
    L.sh_stream_create(ctypes.byref(st))
    out = {"form": "les_speed", "grid": n, "steps": steps, "reps": reps, "clip": clip}
    for name in ("model", "model_lagged", "no_model"):
        m = None
        if name != "no_model":
            m = closure.MLP(seed=3, strm=st)
            m.set_scaling([1.0] * 9, [1e-3] * 6)
        les = closure.LES(n, nu, dt, m, clip=clip, strm=st, lag=int(name == "model_lagged"))
        torch.cuda.synchronize()
        les.set_initial(ub)
        if m is not None:
            les.autotune(2)
        les.steps(2)
        L.sh_stream_synchronize(st)
        L.sh_graph_begin(st)
        les.steps(steps)
        g = ctypes.c_void_p()
        L.sh_graph_end(st, ctypes.byref(g))
        L.sh_graph_launch(g, st)
        L.sh_stream_synchronize(st)
        t0 = time.perf_counter()
        for _ in range(reps):
            L.sh_graph_launch(g, st)
        L.sh_stream_synchronize(st)
        out[name + "_steps_per_s"] = reps * steps / (time.perf_counter() - t0)
        L.sh_graph_destroy(g)
        les.close()
        if m:
            m.close()
    L.sh_stream_destroy(st)
    return out


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="form", required=True)
    q = sub.add_parser("les-speed")
    q.add_argument("--n", type=int, default=64)
    q.add_argument("--steps", type=int, default=200)
    q.add_argument("--reps", type=int, default=5)
    f = sub.add_parser("fp32")
    f.add_argument("--n", type=int, default=512)
    f.add_argument("--steps", type=int, default=100)
    f.add_argument("--dt", type=float, default=2.5e-4)
    f.add_argument("--nu", type=float, default=1.0 / 800)
    f.add_argument("--seed", type=int, default=2024)
    e = sub.add_parser("ensemble")
    e.add_argument("--steps", type=int, default=500)
    e.add_argument("--every", type=int, default=10)
    e.add_argument("--ratio", type=int, default=4)
    e.add_argument("--dt", type=float, default=1e-3)
    e.add_argument("--seed", type=int, default=4242)
    e.add_argument("--out", default=None)
    e.add_argument("--fp64", action="store_true")
    s = sub.add_parser("les")
    s.add_argument("--n-dns", type=int, default=256)
    s.add_argument("--ratio", type=int, default=4)
    s.add_argument("--nu", type=float, default=1.0 / 800)
    s.add_argument("--dt-dns", type=float, default=5e-4)
    s.add_argument("--dt-les", type=float, default=2e-3)
    s.add_argument("--t-end", type=float, default=2.0)
    s.add_argument("--snap-dt", type=float, default=0.25)
    s.add_argument("--eval-dt", type=float, default=0.5)
    s.add_argument("--train-seeds", type=int, default=8)
    s.add_argument("--test-seeds", type=int, default=4)
    s.add_argument("--iters", type=int, default=3000)
    s.add_argument("--batch", type=int, default=65536)
    s.add_argument("--lr", type=float, default=2e-3)
    s.add_argument("--force-iters", type=int, default=2000)
    s.add_argument("--force-lr", type=float, default=1e-3)
    s.add_argument("--select", choices=("mean", "max"), default="mean",
                   help="calibration score: mean or maximum relative E(k) error over k")
    s.add_argument("--variants", default="11,01,12,02",
                   help="comma-separated (clip bit, dealias mode) pairs tried on the training seeds")
    a = p.parse_args(argv)
    torch = _torch()
    torch.cuda.set_device(0)
    lib().sh_init(0)
    if a.form == "fp32":
        print(json.dumps(fp32_error(a.n, a.steps, a.dt, a.nu, a.seed)), flush=True)
    elif a.form == "ensemble":
        print(json.dumps(ensemble_run(steps=a.steps, every=a.every, ratio=a.ratio, dt=a.dt,
                                      precision=SH_F64 if a.fp64 else SH_F32, seed=a.seed,
                                      out=a.out)), flush=True)
    elif a.form == "les-speed":
        print(json.dumps(les_speed(a.n, a.steps, a.reps)), flush=True)
    elif a.form == "les":
        print(json.dumps(les_study(a.n_dns, a.ratio, a.nu, a.dt_dns, a.dt_les, a.t_end, a.snap_dt,
                                   a.train_seeds, a.test_seeds, iters=a.iters, batch=a.batch, lr=a.lr,
                                   force_iters=a.force_iters, force_lr=a.force_lr,
                                   variants=[(v[0] == "1", int(v[1])) for v in a.variants.split(",")],
                                   eval_dt=a.eval_dt, select=a.select)), flush=True)


if __name__ == "__main__":
    sys.exit(main())

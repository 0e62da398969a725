"""What the trained closure gains over the coarse solve without it.

A reduced run of the learned-closure study (spectral_upcxx_amd.forms.les_study):
FP64 DNS at 128^3 (the config-2 IC family: random solenoidal, k0 = 4,
u_rms = 1, nu = 1/800), box-filtered by 4 to 32^3; the network is trained on
the GPU by the project's own kernels on two training seeds (stress fit, then
force matching through the spectral divergence); its gain, dealiasing mode
and clip are chosen on those training seeds; then ONE held-out seed is run
to t = 0.5 by the FP32 32^3 LES with the trained closure and by the same
coarse solver without it (the fallback), both judged against the filtered
DNS shell spectrum E(k), k <= 15, with the deterministic, oracle-pinned
spectrum (sh_spectrum_z).  The closure must cut the held-out error of the
no-model solve by at least a third.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_trained_closure_beats_no_model(hip):
    from spectral_upcxx_amd import forms
    r = forms.les_study(n_dns=128, ratio=4, t_end=0.5, snap_dt=0.05, eval_dt=0.25, train_seeds=2,
                        test_seeds=1, seed0=9100, iters=600, batch=16384, force_iters=300,
                        variants=((True, 1), (False, 2)), gain_candidates=(0.5, 1.0, 1.5),
                        force_variants=(("force", False, 0.0, 1),), cs_candidates=(0.0, 0.05), kcut=16)
    key = "max_rel_err_k_le_15"
    ev = r["evals"][0]
    for name in ("model", "no_model", "no_model_plain"):
        print(name, [(round(x["t"], 3), round(x[key], 4), round(x["mean_rel_err_k_le_15"], 4)) for x in ev[name]],
              flush=True)
    print("chosen network %s clip %s dealias %s gain %s Cs %s" % (r["network"], r["clip"], r["dealias"],
                                                                   r["best_gain"], r["smagorinsky_cs"]), flush=True)
    assert r["dns_all_finite"]
    model, plain_dealiased = r["mean_rel_err_model"], r["mean_rel_err_no_model"]
    print("held-out mean E(k) error: closure %.4f, no model (same dealiasing) %.4f, no model (reference "
          "scheme) %.4f" % (model, plain_dealiased, r["mean_rel_err_no_model_plain"]), flush=True)
    assert math.isfinite(model)
    assert model < (2.0 / 3.0) * plain_dealiased, (model, plain_dealiased)
    assert model < r["mean_rel_err_no_model_plain"]
    # at the last time as well, the closure's worst shell beats the fallback's
    assert ev["model"][-1][key] < ev["no_model"][-1][key]

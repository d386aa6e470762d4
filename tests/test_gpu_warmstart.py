"""What the trained network gains over the method without it, end to end on
a small scale: the batched engine generates steady 64x64 obstacle cavities,
the BF16-MFMA network is trained on them by the engine's own kernels, and
held-out cavities are run to the 1e-8 steady criterion from rest (cold), from
the network's prediction (warm, mfa_ensemble_init_warm; warm_neq, with the
Chapman-Enskog non-equilibrium part of the prediction added) and from the
cold start's own converged fields (exact, exact_neq).  The warm starts must
need fewer steps in total and land on the cold start's fields; from exact
fields the non-equilibrium part must save steps over feq alone."""
import math

import numpy as np
import pytest

from muphfasa_amd import ensemble as E
from muphfasa_amd import surrogate as S
from muphfasa_amd import warmstart as W

pytestmark = pytest.mark.gpu

N = 64
U = 0.1


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    masks = np.stack([S.random_obstacles(rng, N, max_boxes=2, side=(4, 12)) for _ in range(n)])
    Re = np.exp(rng.uniform(math.log(100.0), math.log(400.0), size=n))
    tau = 0.5 + 3.0 * U * N / Re
    return masks, Re, tau


def test_trained_network_shortens_the_run_to_steady_state(dev):
    import torch
    n_train, n_eval = 128, 24
    cases = _cases(n_train, 11)
    masks, Re, ens, r, stats = W.generate(n_train, 0, dtype="f64", tol=1e-7, max_steps=80000,
                                          N=N, cases=cases)
    print("generate", stats, flush=True)
    targets = W.device_targets(ens, n_train, N=N)
    del ens
    keep = (r["status"] == "converged")
    assert keep.sum() > 0.8 * n_train
    idx = torch.from_numpy(np.nonzero(keep)[0]).to(dev)
    targets = targets.index_select(0, idx).contiguous()
    masks_dev = torch.from_numpy(masks[keep]).to(dev)
    re_dev = torch.from_numpy(S.re_feature(Re[keep])).to(dev)
    net = S.WarmStartNet(H=N, W=N, seed=3)
    hist, t_train = W.train(net, masks_dev, re_dev, targets, epochs=400, batch=16, lr=2e-3,
                            seed=0, seconds=45)
    print(f"train: {len(hist)} samples of the loss, first {hist[0][1]:.3e}, last {hist[-1][1]:.3e}, "
          f"{t_train:.1f} s", flush=True)
    assert hist[-1][1] < 0.05 * hist[0][1]
    summary, per_case = W.evaluate(net, n_eval, tol=1e-8, max_steps=150000, N=N,
                                   modes=("cold", "warm", "warm_neq", "exact", "exact_neq"),
                                   cases=_cases(n_eval, 99))
    print({k: v for k, v in summary.items() if k != "bands"}, flush=True)
    assert summary["cold_converged"] >= n_eval - 1
    assert summary["warm_both_converged"] >= n_eval - 1
    # the gain: fewer steps in total, and more than 1 case in 2 faster
    both = (per_case["status_warm"] == "converged") & (per_case["status_cold"] == "converged")
    sw, sc = per_case["steps_warm"][both], per_case["steps_cold"][both]
    assert sw.sum() < sc.sum()
    assert (sw < sc).mean() > 0.5
    # and the same converged fields (max |u_warm - u_cold| / U)
    assert summary["warm_field_diff_max"] < 1e-6
    assert summary["pred_err_l2_median"] < 0.2
    # the Chapman-Enskog start of the prediction: the same converged fields,
    # and fewer steps than the cold start
    both_n = (per_case["status_warm_neq"] == "converged") & (per_case["status_cold"] == "converged")
    assert both_n.sum() >= n_eval - 1
    assert per_case["steps_warm_neq"][both_n].sum() < per_case["steps_cold"][both_n].sum()
    assert summary["warm_neq_field_diff_max"] < 1e-6
    # from the exact fields the non-equilibrium part saves steps over feq alone
    both_x = ((per_case["status_exact"] == "converged") & (per_case["status_exact_neq"] == "converged")
              & (per_case["status_cold"] == "converged"))
    assert per_case["steps_exact_neq"][both_x].sum() < per_case["steps_exact"][both_x].sum()
    assert per_case["steps_exact"][both_x].sum() < per_case["steps_cold"][both_x].sum()
    assert summary["exact_neq_field_diff_max"] < 1e-6

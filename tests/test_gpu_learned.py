"""Learned nonlinear step on gfx950: kernels vs the CPU oracle, training, and the march.

Tolerances: the kernels store activations and back-propagated gradients in
bfloat16 exactly where the oracle rounds (oracle/net_oracle.py, quant=True),
but accumulate in float32 in another order, so a few activations land on the
other side of a bfloat16 rounding boundary (one bf16 ulp, 2^-8 relative).
Forward outputs are compared at rel-L2 <= 5e-3 and gradients at <= 2e-2.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import net_oracle as NO

pytestmark = pytest.mark.gpu


def _net(nT, max_cols=64, seed=0, perturb=0.0):
    from angularspectrum_heterogeneous_amd.learned import LearnedStep, init_params
    p = init_params(seed)
    if perturb:
        rng = np.random.default_rng(seed + 100)
        p = (p + perturb * rng.normal(size=p.shape).astype(np.float32) * NO.trainable_mask()).astype(np.float32)
    return LearnedStep(nT, max_cols=max_cols, params=p), p


def _columns(n, nT, seed):
    from angularspectrum_heterogeneous_amd.learned import tone_bursts
    rng = np.random.default_rng(seed)
    return tone_bursts(n, nT, 4e-8, 1e6, rng)


@pytest.mark.parametrize("nT", [256, 1024])
def test_forward_matches_oracle(nT):
    print(f"[learned] forward nT={nT}", flush=True)
    net, p = _net(nT, perturb=0.05)
    u = _columns(6, nT, 1)
    y, amp = net.forward(torch.from_numpy(u).cuda())
    X0, A = NO.features(u)
    _, yo = NO.forward(p, X0)
    assert np.array_equal(amp.cpu().numpy(), A)
    assert rel_l2(y.cpu().numpy(), yo) <= 5e-3


def test_gradients_match_oracle():
    print("[learned] gradients nT=256", flush=True)
    nT = 256
    net, p = _net(nT, perturb=0.05, seed=2)
    u = _columns(4, nT, 3)
    un = (u * 1.001 + 50.0 * np.roll(u, 1, axis=1) / 1e5).astype(np.float32)
    c = 3.7e-9
    loss, g = net.loss_grad(torch.from_numpy(u).cuda(), torch.from_numpy(un).cuda(), c)
    g = g.cpu().numpy()
    X0, A = NO.features(u)
    acts, yo = NO.forward(p, X0)
    ys = (un.astype(np.float64) - u) / (c * A.astype(np.float64)[:, None] ** 2)
    lo, go = NO.loss_and_grads(p, acts, yo, ys, np.ones_like(yo))
    assert abs(loss - lo) <= 1e-3 * lo
    for j, ((gw, gb), (ow, ob)) in enumerate(zip(NO.unpack(g), NO.unpack(go))):
        assert rel_l2(gw, ow) <= 2e-2, f"layer {j + 1} weights"
        if np.linalg.norm(ob) > 0:
            assert rel_l2(gb, ob) <= 2e-2, f"layer {j + 1} bias"


@pytest.mark.parametrize("nT", [256, 1024])
def test_fused_apply_matches_oracle(nT):
    print(f"[learned] fused apply nT={nT}", flush=True)
    net, p = _net(nT, perturb=0.05, seed=4)
    u = _columns(40, nT, 5)
    # make a few columns steep (square-ish waves) so that they fall back to KT
    u[:5] = (np.sign(u[:5]) * np.abs(u[:5]) ** 0.3 * np.max(np.abs(u[:5]), axis=1, keepdims=True) ** 0.7)
    # and a few weak enough that one step changes them by less than 1e-5 of their amplitude
    u[5:9] *= 1e-3
    u = u.astype(np.float32)
    c = 3.7e-9
    weak = np.float32(c) * np.max(np.abs(np.diff(u, axis=1)), axis=1) < np.float32(1e-5)
    assert weak[5:9].all() and not weak[9:].any()
    # threshold in the middle of a wide gap between the columns' max|du/dt|
    d = np.sort(NO.max_dpdt(u, 4e-8))
    k = max(range(len(d) // 2, len(d) - 1), key=lambda i: d[i + 1] / d[i])
    thr = 0.5 * (d[k] + d[k + 1])
    ud = torch.from_numpy(u).cuda()
    flags = net.apply(ud, c, thr, 4e-8).cpu().numpy()
    want_flag = NO.max_dpdt(u, 4e-8) > thr
    assert np.array_equal(flags.astype(bool), want_flag)
    assert want_flag.any() and not want_flag.all()
    got = ud.cpu().numpy()
    ref = NO.apply(p, u, c)
    keep = ~want_flag
    assert np.array_equal(got[want_flag], u[want_flag])
    assert np.array_equal(got[weak], u[weak]) and np.array_equal(ref[weak], u[weak])
    assert rel_l2(got[keep] - u[keep], ref[keep] - u[keep]) <= 5e-3


def test_learned_path_all_kt_equals_plain():
    """With every column over the max|dp/dt| threshold the learned path is the KT march
    through the real scratch field: identical to the fused KT column pass up to round-off."""
    print("[learned] all-KT plumbing, config 1", flush=True)
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.learned import LearnedStep
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    field0, kw = W.config1()
    net = LearnedStep(field0.shape[2], max_cols=16)
    ref = angular_spectrum_solve(field0, SolverParams(**kw), verbose=False)
    got = angular_spectrum_solve(field0, SolverParams(**kw, ash_learned=net, ash_dpdt_max=0.0), verbose=False)
    for i, name in enumerate(("field", "pnp", "ppp", "pI", "pIloss")):
        assert rel_l2(got[i], ref[i]) <= 1e-6, name


def _train_small(nT, n_pairs, epochs, seed=0):
    from angularspectrum_heterogeneous_amd import learned as L
    from angularspectrum_heterogeneous_amd.solver import SolverParams
    from angularspectrum_heterogeneous_amd import workloads as W
    _, kw = W.config1()
    sp = SolverParams(**kw)
    dz = sp.dZmin
    c = dz * L.burgers_coefficient(sp) / sp.dT
    before, after = L.generate_pairs(sp, nT, dz, n_pairs=n_pairs, stride=8, pairs_per_column=4, seed=seed)
    net = L.LearnedStep(nT, max_cols=256, seed=seed)
    return L, sp, dz, c, before, after, net


def test_training_gain_over_untrained_and_linear():
    """Training on KT pairs: the trained net reproduces the KT increment far better than
    the untrained net, and its one-step error is far below dropping the nonlinear update."""
    print("[learned] training nT=256", flush=True)
    nT = 256
    L, sp, dz, c, before, after, net = _train_small(nT, 32768, 6)
    held_b, held_a = L.generate_pairs(sp, nT, dz, n_pairs=2048, stride=8, pairs_per_column=4, seed=99)

    def inc_err():
        u = held_b.clone()
        net.apply(u, c, 1e30, sp.dT)
        return float(torch.linalg.norm(u - held_a) / torch.linalg.norm(held_a - held_b))

    e0 = inc_err()
    hist = L.train(net, before, after, c, epochs=6, batch=256, lr=2e-3)
    e1 = inc_err()
    print(f"  loss {hist[0]:.3e} -> {hist[-1]:.3e}; increment error untrained {e0:.3e} trained {e1:.3e}")
    assert hist[-1] < hist[0] / 20
    assert e1 < 0.05 and e1 < e0 / 20
    # dropping the update leaves the whole increment as error (ratio 1)


@pytest.mark.timeout(600)
def test_learned_march_config1():
    """Config-1 march with the trained network vs plain KT (rel-L2 <= 1e-3 on the field)."""
    print("[learned] config-1 march", flush=True)
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    nT = 256
    L, sp, dz, c, before, after, net = _train_small(nT, 65536, 8)
    L.train(net, before, after, c, epochs=8, batch=256, lr=2e-3)
    field0, kw = W.config1()
    ref = angular_spectrum_solve(field0, SolverParams(**kw), verbose=False)
    got = angular_spectrum_solve(field0, SolverParams(**kw, ash_learned=net, ash_dpdt_max=3e12), verbose=False)
    lin = angular_spectrum_solve(field0, SolverParams(**{**kw, "beta": 0.0}), verbose=False)
    names = ("field", "pnp", "ppp", "pI", "pIloss")
    errs = {n: rel_l2(got[i], ref[i]) for i, n in enumerate(names)}
    lin_err = rel_l2(lin[0], ref[0])
    print(f"  learned vs KT {errs}; linear (no nonlinear update) vs KT field {lin_err:.2e}")
    assert errs["field"] <= 1e-3
    assert errs["field"] < lin_err / 10

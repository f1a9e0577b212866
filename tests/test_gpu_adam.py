"""The fused HIP Adam step (dtv_adam_step / dtv_adam_step_device through optim.FusedAdam) against torch.optim.Adam
and against the oracle's restatement of it (pytests/train.py:318,351-356: Adam, then the colour clamp
c -= min(c, 0)).

  * golden trajectories written by torch.optim.Adam (tests/golden/adam_golden.npz, 12 steps over gradients
    spanning four decades, colours clamped): every step within 2e-6 relative (torch's CPU kernels order the
    v-update product differently: the oracle meets the same bar);
  * the benchmark's sizes (positions 3V = 823,875 values, not a multiple of 4, so the scalar tail runs; colours
    4V = 1,098,500 with the clamp), three steps: bit-identical to the oracle (same operation order, IEEE
    division and square root) with the step count on the host, and within one float ulp with the step count in
    device memory (the bias corrections' double-precision pow runs on the device then);
  * zero_grad_after leaves zero gradients; the device step counter advances once per step.
"""
import os

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

V_BENCH = 274_625


def _fused(params, lrs, clamps, device_step, zero_grad_after=False):
    from difftetvr_amd.optim import FusedAdam
    groups = [{"params": [p], "lr": lr, "clamp_nonneg": c} for p, lr, c in zip(params, lrs, clamps)]
    return FusedAdam(groups, zero_grad_after=zero_grad_after, device_step=device_step)


@pytest.mark.parametrize("device_step", [False, True])
def test_adam_golden_trajectories(gpu, device_step):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "adam_golden.npz"))
    cases = ("pos", "col")
    params = [torch.from_numpy(g[f"{c}_p0"].copy()).cuda() for c in cases]
    for p in params:
        p.grad = torch.zeros_like(p)
    opt = _fused(params, [float(g[f"{c}_lr"]) for c in cases], [bool(g[f"{c}_clamp"]) for c in cases], device_step)
    nsteps = g["pos_grads"].shape[0]
    worst = 0.0
    for step in range(nsteps):
        for p, c in zip(params, cases):
            p.grad.copy_(torch.from_numpy(g[f"{c}_grads"][step]))
        opt.step()
        for p, c in zip(params, cases):
            ref = g[f"{c}_p"][step]
            got = p.cpu().numpy()
            worst = max(worst, float((np.abs(got - ref) / (np.abs(ref) + 1e-7)).max()))
            assert np.allclose(got, ref, rtol=2e-6, atol=1e-7), (c, step, float(np.abs(got - ref).max()))
    print(f"FusedAdam vs torch.optim.Adam golden (device_step={device_step}): worst rel {worst:.2e}", flush=True)


@pytest.mark.parametrize("device_step", [False, True])
def test_adam_benchmark_sizes_vs_oracle(gpu, device_step):
    rng = np.random.default_rng(11)
    sizes = (3 * V_BENCH, 4 * V_BENCH)
    lrs, clamps = (1e-4, 1e-3), (False, True)
    host = [rng.normal(scale=0.3, size=n).astype(np.float32) for n in sizes]
    host[1] = np.abs(host[1]) * np.float32(1e-3)  # colours near 0: the clamp acts
    params = [torch.from_numpy(h.copy()).cuda() for h in host]
    for p in params:
        p.grad = torch.zeros_like(p)
    opt = _fused(params, lrs, clamps, device_step, zero_grad_after=True)
    om = [np.zeros_like(h) for h in host]
    ov = [np.zeros_like(h) for h in host]
    op = [h.copy() for h in host]
    for step in range(1, 4):
        grads = [(rng.normal(size=n) * 10.0 ** rng.uniform(-3, 1, size=n)).astype(np.float32) for n in sizes]
        for p, gr in zip(params, grads):
            p.grad.copy_(torch.from_numpy(gr))
        opt.step()
        torch.cuda.synchronize()
        for k in range(2):
            oracle.adam(op[k], grads[k], om[k], ov[k], lrs[k], step, clamp_nonneg=clamps[k])
            got = params[k].cpu().numpy()
            st = opt.state[params[k]]
            m, v = st["exp_avg"].cpu().numpy(), st["exp_avg_sq"].cpu().numpy()
            assert np.array_equal(m, om[k]) and np.array_equal(v, ov[k]), (step, k)
            if device_step:
                ulp = np.spacing(np.abs(op[k]).astype(np.float32))
                assert (np.abs(got - op[k]) <= ulp).all(), (step, k, float(np.abs(got - op[k]).max()))
            else:
                assert np.array_equal(got, op[k]), (step, k, int((got != op[k]).sum()))
            assert not params[k].grad.any(), "zero_grad_after left a nonzero gradient"
            if clamps[k]:
                assert (got >= 0).all()
        # the scalar tail (the last 3V mod 4 positions) moved like the rest
        assert not np.array_equal(params[0][-3:].cpu().numpy(), host[0][-3:])
    if device_step:
        assert int(opt._step_state[0].item()) == 3
    print(f"FusedAdam at 3V={sizes[0]}, 4V={sizes[1]} (device_step={device_step}): equal to the oracle", flush=True)

"""Writes adam_golden.npz: parameter trajectories of torch.optim.Adam (the optimiser pytests/train.py:318 uses)
with the colour clamp of train.py:352-356, on seeded inputs.  Run: python tests/golden/make_adam_golden.py"""
import os

import numpy as np
import torch


def trajectory(p0, grads, lr, clamp):
    p = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([p], lr=lr)
    out = []
    for g in grads:
        opt.zero_grad(set_to_none=False)
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        if clamp:
            with torch.no_grad():
                p -= torch.min(p, torch.zeros_like(p))
        out.append(p.detach().numpy().copy())
    return np.stack(out)


def main():
    rng = np.random.default_rng(2024)
    data = {}
    for case, shape, lr, clamp in (("pos", (37, 3), 1e-4, False), ("col", (37, 4), 1e-3, True)):
        p0 = rng.normal(scale=0.5, size=shape).astype(np.float32)
        if clamp:
            p0 = np.abs(p0) * 0.01
        grads = (rng.normal(size=(12,) + shape) * np.logspace(-3, 1, 12)[:, None, None]).astype(np.float32)
        data[f"{case}_p0"] = p0
        data[f"{case}_grads"] = grads
        data[f"{case}_lr"] = np.float64(lr)
        data[f"{case}_clamp"] = np.bool_(clamp)
        data[f"{case}_p"] = trajectory(p0, grads, lr, clamp)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "adam_golden.npz"), **data)


if __name__ == "__main__":
    main()

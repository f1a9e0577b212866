"""dtv_render_noalloc (the allocation-free forward render, through the C ABI) against dtv_render: bit-identical
images on two calls in a row, the second with a smaller mesh and viewport (the shared scratch keeps the first
call's contents), the record-overflow flag, and the size limits."""
import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import grid_case, make_renderer

pytestmark = pytest.mark.gpu


def _noalloc(lib, ctx, r, out, flag):
    desc = r.get_tet_mesh()._desc()
    cam, prm = r._camera(), r._params()
    return lib.dtv_render_noalloc(ctx, ctypes.byref(desc), ctypes.byref(cam), ctypes.byref(prm),
                                  ctypes.c_void_p(out.data_ptr()),
                                  ctypes.c_void_p(flag.data_ptr()) if flag is not None else ctypes.c_void_p(0),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_equals_render_twice(gpu):
    import difftetvr_amd as d
    from difftetvr_amd import _lib
    lib = _lib.load()
    ctx = d._ctx()
    cases = [(grid_case(n=14, seed=31), 160, 120, (0.4, 0.5, 1.8), 3.0, 0),
             (grid_case(n=6, seed=32), 40, 24, (-0.7, 0.2, 1.7), 8.0, 1)]
    for k, (w, W, H, eye, att, order) in enumerate(cases):
        _, r = make_renderer(d, w["cells"], w["pos"], w["col"], W, H, eye=eye, attenuation=att)
        r.set_tile_order("cost" if order == 0 else "row-major")
        ref = r.render()
        out = torch.full((H, W, 4), float("nan"), device="cuda")
        flag = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        rc = _noalloc(lib, ctx, r, out, flag)
        assert rc == 0, lib.dtv_last_error()
        torch.cuda.synchronize()
        print(f"call {k}: {w['cells'].shape[0]} tets, {W}x{H}: overflow flag {int(flag.item())}, "
              f"max |diff| {float((out - ref).abs().max()):.3e}", flush=True)
        assert int(flag.item()) == 0
        assert torch.equal(out, ref)


def test_noalloc_limits_and_overflow(gpu):
    import difftetvr_amd as d
    from difftetvr_amd import _lib
    lib = _lib.load()
    ctx = d._ctx()
    # beyond the pixel limit: refused before any work
    w = grid_case(n=4, seed=1)
    _, r = make_renderer(d, w["cells"], w["pos"], w["col"], 1024, 512)
    out = torch.empty((512, 1024, 4), device="cuda")
    assert _noalloc(lib, ctx, r, out, None) == -1
    assert b"allocation-free limits" in lib.dtv_last_error()
    # more records than the limit: the flag says so (and nothing outside the scratch is touched)
    w = grid_case(n=32, seed=2)   # 196,608 tets
    _, r = make_renderer(d, w["cells"], w["pos"], w["col"], 512, 512, eye=(0.0, 0.0, 1.2))
    r.render()
    k = r.get_stats()["num_pairs"]
    out = torch.empty((512, 512, 4), device="cuda")
    flag = torch.zeros((1,), dtype=torch.int32, device="cuda")
    assert _noalloc(lib, ctx, r, out, flag) == 0, lib.dtv_last_error()
    torch.cuda.synchronize()
    print(f"{k} records against the limit of 786,432: overflow flag {int(flag.item())}", flush=True)
    assert int(flag.item()) == (1 if k > 786_432 else 0)

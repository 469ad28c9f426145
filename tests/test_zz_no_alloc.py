"""rbgpu_trace_noalloc (the allocation-free trace) against rbgpu_trace
through the C ABI: bit-identical rays on two calls in a row, the second
smaller, for the coherence order and the persistent schedule, which are
the parts that work in the library's scratch array; and the limits it
refuses with a message."""
import numpy as np
import pytest

from robast_amd import _abi as A
from robast_amd import configs
from robast_amd._lib import NOALLOC_MAX_RAYS, RbError, make_rays, make_rays_device

pytestmark = pytest.mark.gpu

FIELDS_F = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
FIELDS_I = ("status", "npoints", "last_node")


def _device_rays(ctx, gen, n):
    import torch
    dev = torch.device("cuda")
    t = {f: torch.empty(n, dtype=torch.float64, device=dev) for f in FIELDS_F}
    for f in FIELDS_I:
        t[f] = torch.empty(n, dtype=torch.int32, device=dev)
    ctx.generate(gen, make_rays_device(t), n)
    t["status"][::11] = A.RB_STOP  # not running: left untouched by both entry points
    return t


@pytest.mark.parametrize("config, sizes, order, schedule", [
    ("c2", (1_200_000, 300_000), A.RB_ORDER_AUTO, A.RB_SCHEDULE_AUTO),
    ("c2", (400_000, 70_001), A.RB_ORDER_COHERENT, A.RB_SCHEDULE_PERSISTENT),
    ("c4", (300_000, 50_000), A.RB_ORDER_COHERENT, A.RB_SCHEDULE_AUTO),
    ("c4", (200_000, 30_000), A.RB_ORDER_INPUT, A.RB_SCHEDULE_PERSISTENT),
])
def test_noalloc_trace_equals_trace(ctx, config, sizes, order, schedule):
    import torch
    mgr = {"c2": configs.c2_scene, "c4": configs.c4_scene}[config]()
    gen = configs.c2_raygen() if config == "c2" else configs.c4_raygen(mgr.c4_dz)
    scene = ctx.upload(mgr.flat())
    opts = mgr.trace_opts(device_ptrs=1, order=order, schedule=schedule)
    for n in sizes:
        print(f"noalloc {config} n={n} order={order} schedule={schedule}", flush=True)
        a = _device_rays(ctx, gen, n)
        b = {k: v.clone() for k, v in a.items()}
        ctx.trace(scene, make_rays_device(a), n, opts)
        ctx.trace_noalloc(scene, make_rays_device(b), n, opts)
        for k in a:
            assert torch.equal(a[k], b[k]), (config, n, k)
        st = torch.bincount(b["status"].long(), minlength=6).cpu().numpy()
        assert st[A.RB_RUN] == 0 and st[A.RB_FOCUS] > 0, st


def test_noalloc_trace_refuses_what_it_cannot_do(ctx):
    import torch
    mgr = configs.c2_scene()
    scene = ctx.upload(mgr.flat())
    h = configs.c1_rays(0.0, nside=8).soa()
    # host arrays: would need staging buffers
    with pytest.raises(RbError, match="device memory"):
        ctx.trace_noalloc(scene, make_rays(h), len(h["x"]), mgr.trace_opts())
    # beyond the scratch array: refused before any device work
    t = {f: torch.empty(1, dtype=torch.float64, device="cuda") for f in FIELDS_F}
    for f in FIELDS_I:
        t[f] = torch.empty(1, dtype=torch.int32, device="cuda")
    with pytest.raises(RbError, match="scratch maximum"):
        ctx.trace_noalloc(scene, make_rays_device(t), NOALLOC_MAX_RAYS + 1, mgr.trace_opts(device_ptrs=1))
    # an empty batch is a no-op
    ctx.trace_noalloc(scene, make_rays_device(t), 0, mgr.trace_opts(device_ptrs=1))
    np.testing.assert_equal(ctx.last_trace_launches(), 0)

"""ct_trace_rays_device_noalloc against ct_trace_rays_device: the same
closest-hit results (ids and t, bit for bit) on two calls in a row, the
second with fewer rays, through the C ABI, on a context of its own."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import DeviceArray, ray_set

pytestmark = pytest.mark.gpu


def _device_rays(o, d):
    n = o.shape[0]
    h_o = np.zeros((n, 4), np.float32)
    h_d = np.zeros((n, 4), np.float32)
    h_o[:, :3], h_d[:, :3] = o, d
    return DeviceArray(h_o), DeviceArray(h_d)


@pytest.mark.parametrize("config,accel", [(2, 1), (1, 0)])
def test_noalloc_equals_allocating_entry_point(ct, orc, config, accel):
    from conftest import get_scene
    scene, arrays, osc = get_scene(ct, orc, config)
    ctx = ct.Context(0)
    try:
        ctx.upload(scene)
        opts, _, _ = scene.frame_opts()
        opts.accel = accel
        opts.traversal = ct.TRAV_EXACT
        for n, seed in ((1 << 20, 500 + config), (1 << 14, 600 + config)):
            print(f"config {config} accel {accel}: {n} rays", flush=True)
            o, d = ray_set(scene, arrays, n, seed)
            d_o, d_d = _device_rays(o, d)
            h_a = DeviceArray(np.full((n, 2), -7.0, np.float32))
            h_b = DeviceArray(np.full((n, 2), -9.0, np.float32))
            ms = ctx.trace_rays_device_noalloc(d_o.addr, d_d.addr, n, h_a.addr, opts)
            assert ms > 0
            ctx.trace_rays_device(d_o.addr, d_d.addr, n, h_b.addr, opts)
            a, b = h_a.download(), h_b.download()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            t_ref, p_ref = osc.trace(o, d, accel, 32, opts.triangle_epsilon, threads=16)
            assert np.array_equal(a[:, 1].view(np.int32), p_ref)
            assert np.array_equal(a[:, 0], t_ref.astype(np.float32))
    finally:
        ctx.close()

"""tpt_trace_closest_noalloc: the closest-hit entry point that allocates nothing (its traversal
overflow stacks live in a fixed per-device scratch) against tpt_trace_closest on the same rays:
bit-identical hits, on two calls in a row, the second with fewer rays, on scenes whose
traversal stack runs past the 24 entries per lane kept in LDS (so the scratch is used)."""
import ctypes as C

import numpy as np
import pytest
import torch

import tinypt_amd as T
from tinypt_amd import scenes

from parity_scenes import PARITY_CASES
from trace_checks import stress_rays

pytestmark = pytest.mark.gpu

LDS_ENTRIES = 24


def _noalloc(ds, rays):
    dev = torch.device("cuda", ds.ctx.device)
    r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32)).to(dev)
    n = r.shape[0]
    t = torch.empty(n, dtype=torch.float32, device=dev)
    kind, inst, prim = (torch.empty(n, dtype=torch.int32, device=dev) for _ in range(3))
    torch.cuda.synchronize(dev)
    lib = T.load_library()
    rc = lib.tpt_trace_closest_noalloc(ds.ctx.h, ds.h, C.c_int64(n), C.c_void_p(r.data_ptr()), C.c_void_p(t.data_ptr()),
                                       C.c_void_p(kind.data_ptr()), C.c_void_p(inst.data_ptr()),
                                       C.c_void_p(prim.data_ptr()))
    assert rc == 0, lib.tpt_last_error().decode()
    return dict(t=t.cpu().numpy(), kind=kind.cpu().numpy(), inst=inst.cpu().numpy(), prim=prim.cpu().numpy())


@pytest.mark.parametrize("name", ["instanced_small", "instanced_4096"])
def test_noalloc_trace_equals_trace(gpu_ctx, name):
    if name == "instanced_small":
        sc, box = PARITY_CASES["instanced_small"][0](), 4.5
    else:
        sc, box = scenes.instanced_scene(32, 32, n_variants=2, n_instances=4096, subdivisions=2, box=20.0,
                                         env_size=(16, 8)), 22.0
    ds = T.DeviceScene(sc, gpu_ctx)
    stack = ds.info()["traversal_stack"]
    assert stack > LDS_ENTRIES, stack  # the overflow part of the stack, in the scratch, is used
    rng = np.random.default_rng(31)
    targets = np.array([inst.xform.t for inst in sc.world.instances], dtype=np.float64)
    for n in (600_000, 70_001):  # the second call is smaller; the scratch keeps the first call's contents
        rays = stress_rays(rng, n, (-box,) * 3, (box,) * 3, targets)
        want = ds.trace_closest(rays)
        got = _noalloc(ds, rays)
        for k in ("kind", "inst", "prim"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"{name} n={n} {k}")
        np.testing.assert_array_equal(got["t"].view(np.int32), want["t"].view(np.int32), err_msg=f"{name} n={n} t")
        print(f"{name}: {n} rays, traversal stack {stack}, hit {float((want['kind'] != 0).mean()):.3f}, identical",
              flush=True)
    # an empty batch is a no-op
    lib = T.load_library()
    assert lib.tpt_trace_closest_noalloc(ds.ctx.h, ds.h, C.c_int64(0), None, None, None, None, None) == 0

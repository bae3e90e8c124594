"""No-allocation diffusion (GPU): cfft_diffuse_noalloc_dev, whose work buffers
live in the library's static device array, against cfft_diffuse_dev (the
plan's own buffers) on the same plan and input, bit for bit.  Two calls in a
row on different plans, the second with the smaller input, so the second
call runs over what the first left in the shared array; then the size and
path checks of the entry point."""
import ctypes as C

import numpy as np
import pytest

import cfft_hip as cf
from cfft_hip import _lib as L
import oracle as O
from conftest import group_setup

pytestmark = pytest.mark.gpu

MAX_BYTES = 16 << 20  # the entry point's stated maximum of work memory


def _case(sg, n, path, prec, seed, batch=1, symmetric=True):
    import torch
    N = (n, n, n)
    ops, ops_s, shift, R, t = group_setup(sg, N)
    plan = cf.DevicePlan(N, ops_s, 0.01, 4.0 * np.eye(3), path, prec, batch)
    Lv = O.auto_L(t)
    if symmetric:
        f0 = np.stack([O.extract_subgrid(O.make_symmetric(R, t, N, seed=seed + b), Lv) for b in range(batch)])
    else:  # input of the right size only (for calls that must be refused)
        f0 = np.random.default_rng(seed).uniform(-1.0, 1.0, (batch,) + plan.M)
    dt = torch.float64 if prec == cf.FP64 else torch.float32
    # the subgrid field in the plan's layout: x fastest, batch entries consecutive
    flat = np.concatenate([f.ravel(order="F") for f in f0])
    x = torch.tensor(flat, dtype=dt, device="cuda")
    return plan, x


def _run_both(plan, x):
    import torch
    a = x.clone()
    b = x.clone()
    torch.cuda.synchronize()
    L.check(L.lib().cfft_diffuse_dev(plan.handle, C.c_void_p(a.data_ptr()), C.c_void_p(0)))
    L.check(L.lib().cfft_diffuse_noalloc_dev(plan.handle, C.c_void_p(b.data_ptr()), C.c_void_p(0)))
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("first,second", [
    ((229, 128, cf.PATH_QS, cf.FP64, 11), (47, 32, cf.PATH_M2, cf.FP64, 12)),
    ((230, 64, cf.PATH_M7, cf.FP64, 13), (229, 32, cf.PATH_QS, cf.MIXED, 14)),
    ((70, 48, cf.PATH_M7Q, cf.FP64, 15), (47, 16, cf.PATH_M2Q, cf.FP64, 16)),
])
def test_noalloc_matches_diffuse_dev(first, second):
    for case in (first, second):
        print(f"no-alloc diffuse: sg={case[0]} N={case[1]}^3 path={case[2]} prec={case[3]}", flush=True)
        plan, x = _case(*case)
        a, b = _run_both(plan, x)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), f"max |diff| {np.max(np.abs(a - b)):.3e}"
        assert not np.array_equal(a, x.cpu().numpy())  # the operator ran


@pytest.mark.parametrize("axis", [0, 1])
def test_noalloc_batched_qs(axis):
    """Path QS, 2 fields at Im-3m 128^3, with the plane-pair chain and with
    the per-axis chain (CFFT_OPT_AXIS_SPLIT)."""
    print(f"no-alloc diffuse: Im-3m 128^3 QS, 2 fields, axis={axis}", flush=True)
    plan, x = _case(229, 128, cf.PATH_QS, cf.FP64, 21, batch=2)
    plan.set_option(cf._lib.OPT_AXIS_SPLIT, axis)
    a, b = _run_both(plan, x)
    assert np.array_equal(a, b)


def test_noalloc_rejects_oversize_and_qr():
    import torch
    # Im-3m 128^3 FP64 QS needs 2 MiB of channels per field: 9 fields exceed 16 MiB
    plan, x = _case(229, 128, cf.PATH_QS, cf.FP64, 31, batch=9, symmetric=False)
    before = x.clone()
    r = L.lib().cfft_diffuse_noalloc_dev(plan.handle, C.c_void_p(x.data_ptr()), C.c_void_p(0))
    assert r == L.CFFT_ERR_SIZE
    assert "maximum" in L.last_error()
    torch.cuda.synchronize()
    assert torch.equal(x, before)  # nothing ran
    plan, x = _case(229, 32, cf.PATH_QR, cf.FP64, 32, symmetric=False)
    r = L.lib().cfft_diffuse_noalloc_dev(plan.handle, C.c_void_p(x.data_ptr()), C.c_void_p(0))
    assert r == L.CFFT_ERR_UNSUPPORTED

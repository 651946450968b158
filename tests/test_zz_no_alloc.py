"""k2_eval_device_noalloc (partial sums in the library's static scratch, no device allocation)
against k2_eval_device, called through the C ABI: two calls in a row, the second smaller, plus
the Z-tile path, the mixed form and the error paths."""
import ctypes

import numpy as np
import pytest

from mertens_bound_amd import evaluator as ev
from mertens_bound_amd import synth

pytestmark = pytest.mark.gpu


def _lib():
    L = ev.load_library()
    L.k2_eval_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                         ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
    return L


def _noalloc(t, Vt, k, flags, torch):
    n = Vt.shape[0]
    out = [torch.full((n,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(3)]
    st = torch.cuda.current_stream().cuda_stream
    rc = _lib().k2_eval_device_noalloc(t.handle, n, Vt.data_ptr(), k, 0, flags,
                                       out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), st)
    return rc, out


@pytest.fixture(scope="module")
def setup():
    import torch
    tb = synth.make_zero_table(100_000, 41)
    t = ev.K2Table(tb)
    # at this tiling 131 splits x 64 candidate tiles would need 206 MB of partial sums for
    # 65,536 candidates, so the allocation-free call cuts the plan to fit 64 MiB
    t.set_tiling(nc=4, target_blocks=12288)
    return tb, t, torch


@pytest.mark.parametrize("flags", [ev.VALUE | ev.D1, ev.VALUE | ev.D1 | ev.D2 | ev.MIXED])
def test_noalloc_matches_eval_device_two_calls(setup, flags):
    tb, t, torch = setup
    k = 1.5e-10
    t.mixed_head(k)                             # makes k's weights resident
    sizes = (65_536, 1_000)                     # the second call smaller
    for i, n in enumerate(sizes):
        V = synth.make_candidates(n, 3.4e18, 3.6e18, 50 + i)
        Vt = ev.v_tensor(V)
        rc, got = _noalloc(t, Vt, k, flags, torch)
        assert rc == 0, ev.load_library().k2_last_error()
        want = t.eval_device(Vt, k, flags=flags)
        torch.cuda.synchronize()
        for q, f in enumerate((ev.VALUE, ev.D1, ev.D2)):
            if not flags & f:
                continue
            g, w = got[q].cpu().numpy(), want[q].cpu().numpy()
            assert np.all(np.isfinite(g))
            scale = np.sum(2 * np.exp(-k * tb.gamma ** 2) * np.abs(tb.w) * tb.gamma ** q)
            err = np.max(np.abs(g - w)) / scale
            print(f"n={n} flags={flags} sum {q}: max diff {err:.2e} of its scale")
            if n == 1_000 and not flags & ev.MIXED:
                # the plan (1,563 splits of 64 zeros, 38 MB) fits the scratch unchanged: the
                # same splits, bit for bit.  The mixed form's head and tail plans need about
                # 3,100 rows (77 MB) at this size and are cut like the large call's.
                assert np.array_equal(g, w)
            else:
                # fewer, longer splits: the same terms summed in another order.  Each order's
                # rounding is within (zeros per split + splits) u sum|a| of the exact sum
                # (about 2,500 u here); the bar is 1e-13 of the scale in FP64, and the
                # precision form's 1e-9 where its FP32 tail sums differ
                assert err <= (1e-13 if not flags & ev.MIXED else 1e-9)


def test_noalloc_few_candidates_and_errors(setup):
    tb, t, torch = setup
    k = 1.5e-10
    t.mixed_head(k)
    V = synth.make_candidates(100, 3.4e18, 3.6e18, 60)      # Z tiles
    Vt = ev.v_tensor(V)
    rc, got = _noalloc(t, Vt, k, ev.VALUE | ev.D1, torch)
    assert rc == 0
    want = t.eval_device(Vt, k, flags=ev.VALUE | ev.D1)
    torch.cuda.synchronize()
    assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])
    # a k whose weights are not resident is refused (computing them would allocate)
    rc, _ = _noalloc(t, Vt, 7.25e-10, ev.VALUE, torch)
    assert rc == 1 and b"not resident" in ev.load_library().k2_last_error()
    # more candidates than the scratch can hold partial sums for (3 x 8 B x 3,000,064 > 64 MiB)
    big = torch.zeros((3_000_000, 2), dtype=torch.int64, device="cuda")
    rc, _ = _noalloc(t, big, k, ev.VALUE, torch)
    assert rc == 1 and b"scratch" in ev.load_library().k2_last_error()

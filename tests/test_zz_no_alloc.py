"""maomi_v4_apply_split_f64_noalloc (partial sums in the library's static device array)
against maomi_v4_apply_split_f64 (partial sums in a caller buffer), through the C ABI.

Bar: bit-identical outputs (the same kernels sum the same terms in the same order), on two
calls in a row, the second with a smaller input, so the second call reuses the array the
first one filled.  The size limit is checked on the host before any GPU work."""

import numpy as np
import pytest

from maomi_amd import _native as nat

LIMIT = 4 * 1024 * 1024  # MAOMI_V4_STATIC_PARTIALS


def test_noalloc_rejects_partials_beyond_the_static_buffer():
    lib = nat.lib()
    one = 8  # a non-null address: the call returns before touching it
    rc = lib.maomi_v4_apply_split_f64_noalloc(one, one, one, one, 10, LIMIT // 6 + 1, 3, 2, None)
    assert rc == -1
    assert b"exceed the static buffer" in lib.maomi_last_error()
    rc = lib.maomi_v4_apply_split_f64_noalloc(None, one, one, one, 10, 5, 1, 2, None)
    assert rc == -1 and b"bad arguments" in lib.maomi_last_error()


@pytest.mark.gpu
def test_noalloc_equals_caller_buffer_entry_point(gpu_lib):
    import torch
    dev = torch.device("cuda")
    st = nat.stream_handle(dev)
    rng = np.random.default_rng(17)
    # the 4225-site lattice on the 512^2 pixel grid (the config-2 mask), then a smaller case
    t = np.round(np.linspace(0, 512, 65))
    xx, yy = np.meshgrid(t, t, indexing="ij")
    big_p = np.stack([xx.ravel(), yy.ravel()], 1)
    g = np.arange(512) + 0.5
    qx, qy = np.meshgrid(g, g, indexing="ij")
    big_q = np.stack([qx.ravel(), qy.ravel()], 1)
    cases = [(big_p, big_q, 3), (rng.random((3000, 2)) * 90, rng.random((2000, 2)) * 100, 2)]
    for p, q, C in cases:
        P, Q = len(p), len(q)
        rows = int(gpu_lib.maomi_v4_apply_rows(P, Q))
        assert rows > 1 and rows * Q * C <= LIMIT
        coef = rng.standard_normal((P, C))
        tp, tc, tq = (torch.as_tensor(a, device=dev) for a in (p, coef, q))
        ref = torch.empty((Q, C), dtype=torch.float64, device=dev)
        part = torch.empty(rows * Q * C, dtype=torch.float64, device=dev)
        print(f"[no-alloc v4 apply] P={P} Q={Q} C={C} rows={rows}", flush=True)
        nat.check(gpu_lib.maomi_v4_apply_split_f64(nat.ptr(tp), nat.ptr(tc), nat.ptr(tq),
                                                   nat.ptr(ref), P, Q, C, nat.ptr(part), rows,
                                                   st), "v4 apply")
        out = torch.full((Q, C), float("nan"), dtype=torch.float64, device=dev)
        nat.check(gpu_lib.maomi_v4_apply_split_f64_noalloc(nat.ptr(tp), nat.ptr(tc), nat.ptr(tq),
                                                           nat.ptr(out), P, Q, C, rows, st),
                  "v4 apply without allocation")
        torch.cuda.synchronize()
        assert torch.equal(out, ref)

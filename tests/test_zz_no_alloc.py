"""sq_tridiag_nodes_weights_noalloc through the C ABI.

The QL nodes and weights of device tridiagonals with the workspace in the
library's static device array, against sq_tridiag_nodes_weights (the same QL
kernel on host buffers with an allocated workspace): bit-identical on two calls
in a row, the second smaller (the array keeps the first call's contents).  A
tridiagonal whose QL cannot converge (NaN entries: the kernel's 50-sweep limit
ends it) is reported by index as the allocating entry point reports it; a
batch beyond the static workspace and host buffers are rejected before any
launch.
"""
import numpy as np
import pytest
import torch

from conftest import MG1

import sqminiapp_amd as sq

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _tridiags(rng, count, n):
    # Lanczos-like: diagonal inside a spectrum's range, positive subdiagonal
    return rng.uniform(-1.0, 25.0, (count, n)), rng.uniform(0.5, 9.0, (count, n))


def _device_call(ctx, d, s):
    count, n = d.shape
    dd = torch.from_numpy(np.ascontiguousarray(d)).to(DEV)
    ds = torch.from_numpy(np.ascontiguousarray(s)).to(DEV)
    out = [torch.full((count, n), float("nan"), dtype=torch.float64, device=DEV) for _ in range(2)]
    out += [torch.full((count,), float("nan"), dtype=torch.float64, device=DEV) for _ in range(2)]
    torch.cuda.synchronize()      # the entry point orders nothing against torch's stream
    ctx.tridiag_nodes_weights_noalloc(dd.data_ptr(), ds.data_ptr(), n, count, *(t.data_ptr() for t in out))
    return [t.cpu().numpy() for t in out]


@pytest.fixture(scope="module")
def ctx():
    c = sq.Context(MG1, "Mg1", phantom_atom=True, device=0)
    yield c
    c.close()


def test_noalloc_matches_allocating_entry_point(ctx):
    rng = np.random.default_rng(7)
    for count, n in ((9000, 80), (1500, 40)):
        d, s = _tridiags(rng, count, n)
        print(f"no-alloc QL: count {count}, n {n}", flush=True)
        got = _device_call(ctx, d, s)
        ref = ctx.tridiag_nodes_weights(d, s)
        for name, g, r in zip(("lambda", "w", "lmin", "lmax"), got, ref):
            assert np.array_equal(g, r), (count, n, name)
        assert np.max(np.abs(got[1].sum(1) - 1.0)) < 1e-12


def test_noalloc_reports_the_failed_tridiagonal(ctx):
    rng = np.random.default_rng(8)
    d, s = _tridiags(rng, 500, 30)
    d[123, 5] = np.nan
    with pytest.raises(sq.SQError, match="tridiagonal 123"):
        ctx.tridiag_nodes_weights(d, s)
    with pytest.raises(sq.SQError, match="tridiagonal 123"):
        _device_call(ctx, d, s)
    # the next call on the same array is clean again
    d2, s2 = _tridiags(rng, 200, 30)
    got = _device_call(ctx, d2, s2)
    ref = ctx.tridiag_nodes_weights(d2, s2)
    assert all(np.array_equal(g, r) for g, r in zip(got, ref))


def test_noalloc_rejects_oversize_and_host_buffers(ctx):
    small = [torch.zeros(80, dtype=torch.float64, device=DEV) for _ in range(6)]
    p = [t.data_ptr() for t in small]
    with pytest.raises(sq.SQError, match="count <= 17439"):
        ctx.tridiag_nodes_weights_noalloc(p[0], p[1], 80, 17440, p[2], p[3], p[4], p[5])
    host = np.zeros(80)
    with pytest.raises(sq.SQError, match="not device memory"):
        ctx.tridiag_nodes_weights_noalloc(host.ctypes.data, p[1], 80, 1, p[2], p[3], p[4], p[5])

"""gatt_block_dot_noalloc, the allocation-free variant of gatt_block_dot (the block
Gram-Schmidt products H = X^T Y of calcInstability.m:389-393).

It runs the same kernels with the same chunking, so its H equals gatt_block_dot's bit
for bit; two calls in a row (the second on a smaller block) reuse the one static
scratch array of the device, and a block whose partial sums exceed that array is
refused before anything is launched.  The torch float64 product is a sanity bound.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu


def test_block_dot_noalloc_matches_block_dot():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd import runtime
    case = cases.periodic_rest(16, 16)
    s = runtime.Solver(case)
    try:
        g = torch.Generator(device="cuda").manual_seed(7)
        for k, p, n in ((32, 32, 5 * 64 ** 3), (5, 3, 10007)):
            print("block product k=%d p=%d n=%d" % (k, p, n), flush=True)
            X = torch.randn(k, n, generator=g, dtype=torch.float64, device="cuda")
            Y = torch.randn(p, n, generator=g, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            want = s.block_dot(X, Y)
            got = s.block_dot(X, Y, noalloc=True)
            np.testing.assert_array_equal(got, want)
            ref = (X @ Y.T).cpu().numpy()
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(n) * np.max(np.abs(ref))
        # 1024 x 1025 partial sums plus the reduced block do not fit the 2^20 doubles
        X = torch.zeros(1024, 8, dtype=torch.float64, device="cuda")
        Y = torch.zeros(1025, 8, dtype=torch.float64, device="cuda")
        with pytest.raises(runtime.GattError, match="allocation-free limit"):
            s.block_dot(X, Y, noalloc=True)
    finally:
        s.close()

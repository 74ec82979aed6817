"""The split-precision radix-16 stage on INT8 matrix cores (nttm_ntt_fwd_mfma): bit-identical to the FP64 forward
transform (itself equal to the O(n^2) DFT, test_gpu_ntt.py) for every prime, several lengths and batches, random and
extreme residues; a whole product formed with it (config 4's path: limbs -> MFMA forward transforms -> pointwise
inverse -> CRT) equals nttm_mul's; the live profile shows the matrix-core kernel ran."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prime", [0, 1, 2, 3])
@pytest.mark.parametrize("logn,batch", [(8, 1), (9, 3), (12, 2), (16, 1), (20, 1)])
def test_fwd_mfma_equals_fwd(nttm, prime, logn, batch):
    p = nttm.prime(prime)
    rng = np.random.default_rng(prime * 100 + logn)
    x = rng.integers(0, p, batch << logn, dtype=np.uint64)
    x[:37] = p - 1                                                # extreme residues: every digit at its maximum
    x[37:50] = 0
    a, b = dev(x), dev(x)
    say("fwd mfma", prime, logn, batch)
    nttm.ntt_fwd(prime, a, logn, batch)
    nttm.ntt_fwd_mfma(prime, b, logn, batch)
    assert np.array_equal(host(a), host(b))


def test_fwd_mfma_matches_dft(nttm):
    """independent of the FP64 engine: 2^8 points against the O(n^2) DFT of the oracle"""
    p = nttm.prime(0)
    x = np.random.default_rng(3).integers(0, p, 256, dtype=np.uint64)
    t = dev(x)
    say("fwd mfma dft")
    nttm.ntt_fwd_mfma(0, t, 8)
    assert np.array_equal(host(t), O.dft_bitrev(x, p, nttm.root(0, 8)))


def test_product_through_mfma_stage(nttm):
    """c = a b for 2^15 x 2^15 limbs with every forward transform through the matrix-core stage: equal to nttm_mul"""
    import torch
    n = 1 << 15
    a, s = O.gen_limbs(n, 0x5EED0003)
    b, _ = O.gen_limbs(n, 0x5EED0004)
    da, db = dev(a), dev(b)
    logn = 16
    planes = []
    say("product through the mfma stage")
    nttm.profile(True)
    for i in range(3):
        pa = torch.empty(1 << logn, dtype=torch.int64, device="cuda")
        pb = torch.empty_like(pa)
        nttm.load(i, pa, da)
        nttm.load(i, pb, db)
        nttm.ntt_fwd_mfma(i, pa, logn)
        nttm.ntt_fwd_mfma(i, pb, logn)
        nttm.ntt_inv_pw(i, pa, pb, 0, logn)
        planes.append(pa)
    stats = nttm.profile_read()
    nttm.profile(False)
    c = nttm.crt(planes, 2 * n - 1, 2 * n)
    want = nttm.mul(da, db)
    assert torch.equal(c, want)
    assert stats["k_dft16_mfma"][0] == 6

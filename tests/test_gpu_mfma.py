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


def test_fwd_mfma_equals_fwd_2e28(nttm):
    """config 4's size: one prime's forward transform of 2^28 points (the 2^27-limb operand, seed 0x5EED0003, loaded
    mod the prime) through the matrix-core stage equals the FP64 transform bit for bit"""
    import torch
    n = 1 << 27
    a, _ = O.gen_limbs(n, 0x5EED0003)
    da = dev(a)
    x = torch.empty(1 << 28, dtype=torch.int64, device="cuda")
    nttm.load(1, x, da)
    del da
    y = x.clone()
    say("fwd mfma 2^28")
    nttm.ntt_fwd(1, x, 28)
    nttm.ntt_fwd_mfma(1, y, 28)
    torch.cuda.synchronize()
    assert torch.equal(x, y)


def test_product_2e27_through_mfma_stage_residue_form(nttm):
    """the 2^27 x 2^27-limb product (seeds 0x5EED0003 / 0x5EED0004) with all six forward transforms through the
    matrix-core stage passes the reference checker's residue form and both 64-limb edges against GMP -- checked
    against the operands, not against the FP64 path"""
    import torch
    n = 1 << 27
    a, _ = O.gen_limbs(n, 0x5EED0003)
    b, _ = O.gen_limbs(n, 0x5EED0004)
    da, db = dev(a), dev(b)
    planes = []
    say("2^27 product through the mfma stage")
    for i in range(3):
        pa = torch.empty(1 << 28, dtype=torch.int64, device="cuda")
        pb = torch.empty_like(pa)
        nttm.load(i, pa, da)
        nttm.load(i, pb, db)
        nttm.ntt_fwd_mfma(i, pa, 28)
        nttm.ntt_fwd_mfma(i, pb, 28)
        nttm.ntt_inv_pw(i, pa, pb, 0, 28)
        del pb
        planes.append(pa)
    del da, db
    c = nttm.crt(planes, 2 * n - 1, 2 * n)
    torch.cuda.synchronize()
    del planes
    ch = host(c)
    del c
    assert O.tref_res(a, b, ch)
    if O.have_gmp():
        assert O.tref_edges(a, b, ch)

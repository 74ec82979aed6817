"""The size sweep's machinery on the GPU: operands generated on the device equal the oracle's generator (the
reference's splitmix64 stream, E/tests/harness.h:55-81); a product checked by device residues mod the eight T1 primes
agrees with GMP where GMP can check it, and the residue check catches a single changed limb; the real-shape share of
one GPU -- 2^31 x 2^31 limbs, 2^32 transform points on four primes, 224 GiB of HBM -- passes the residue check."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


def residues_ok(nttm, da, db, dc):
    ra, rb, rc = nttm.limbs_mod(da), nttm.limbs_mod(db), nttm.limbs_mod(dc)
    return all((x * y) % q == z for x, y, z, q in zip(ra, rb, rc, nttm.T1_PRIMES))


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 0x5EED0003), (1000, 77), (100003, 0x5EED0004)])
def test_device_generator_equals_oracle(nttm, n, seed):
    say("gen", n, seed)
    got = host(nttm.gen_limbs_dev(n, seed))
    want, _ = O.gen_limbs(n, seed)
    assert np.array_equal(got, want)


def test_residue_check_agrees_with_gmp(nttm):
    import torch
    n = 1 << 18
    say("residue check 2^18")
    da, db = nttm.gen_limbs_dev(n, 0x5EED0003), nttm.gen_limbs_dev(n, 0x5EED0004)
    dc = nttm.mul(da, db)
    assert residues_ok(nttm, da, db, dc)
    if O.have_gmp():
        assert O.gmp_check(host(da), host(db), host(dc))
    bad = dc.clone()
    bad[12345] += 1
    assert not residues_ok(nttm, da, db, bad)
    torch.cuda.synchronize()


def test_real_shape_share_2e31_limbs(nttm):
    """2^31 x 2^31 limbs (transform 2^32, four primes: the product exceeds the three-prime bound of 1,426,725,066
    terms): (np + 1) planes of 32 GiB and 64 GiB of operands and result fit one GPU; every limb below 10^18, the
    normalized length 2^32 or 2^32 - 1, and c = a b modulo the eight T1 primes"""
    import torch
    n = 1 << 31
    free, total = torch.cuda.mem_get_info()
    if free < (236 << 30):
        pytest.skip("needs about 236 GiB of free HBM (%.0f GiB free)" % (free / 2 ** 30))
    say("2^31 x 2^31 limbs")
    da, db = nttm.gen_limbs_dev(n, 0x5EED0003), nttm.gen_limbs_dev(n, 0x5EED0004)
    dc = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    nttm.mul(da, db, out=dc)
    torch.cuda.synchronize()
    assert bool(((dc >= 0) & (dc < O.B)).all().item())
    assert int(dc[-1].item()) != 0 or int(dc[-2].item()) != 0
    assert residues_ok(nttm, da, db, dc)
    del da, db, dc
    torch.cuda.empty_cache()
    nttm.shutdown(); nttm.init(0)                                  # give back the 160 GiB workspace

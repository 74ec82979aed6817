"""The residue post-condition (nttm_config.check, nttm_check_product): c = a b + x modulo the prime 2^62 + 135 ... =
4611686018427388039, the identity the reference's "detect and retry" checks on its tracked residues (E/dr.h:1-8,
ALGORITHM.md:441-458).  With the check on, correct products of every entry pass; on demand it accepts the correct
result and rejects one with a single limb changed (low, middle, top), a changed x and a truncated result -- a wrong
value, reported as NTTM_ECHECK, never a fault; its residues agree with the oracle's on the host."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
Q = 4611686018427388039


@pytest.fixture
def checked(nttm):
    import torch
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0, check=1)
    yield nttm
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0)


def test_products_pass_with_the_check_on(checked):
    import torch
    a, s = O.gen_limbs(1 << 18, 0x5EED0401); b, s = O.gen_limbs(77777, s); x, _ = O.gen_limbs(300000, s)
    say("checked single products")
    c = checked.mul(dev(a), dev(b))
    c2 = checked.mul_add(dev(a), dev(b), dev(x))
    if O.have_gmp():
        assert O.gmp_check(a, b, host(c)) and O.gmp_check(a, b, host(c2), x)
    ops, s = [], 9
    for i in range(40):
        na, nb = 1 + 97 * i, 5000 - 101 * i
        aa, s = O.gen_limbs(na, s); bb, s = O.gen_limbs(nb, s)
        xx = O.gen_limbs(min(na + nb, 50 + 13 * i), s)[0] if i % 3 == 0 else None
        ops.append((aa, bb, xx))
    items = [(dev(aa), dev(bb), None if xx is None else dev(xx),
              torch.zeros(len(aa) + len(bb) + (1 if xx is not None else 0), dtype=torch.int64, device="cuda"))
             for aa, bb, xx in ops]
    say("checked batch")
    checked.mul_batch(checked.make_prods(items))
    checked.mul_each_async(checked.make_prods(items))
    for (aa, bb, xx), it in zip(ops[:6], items):
        assert np.array_equal(host(it[3]), O.mul_school(aa, bb, xx))


def test_check_rejects_wrong_results(nttm):
    a, s = O.gen_limbs(40000, 0x5EED0402); b, s = O.gen_limbs(30001, s); x, _ = O.gen_limbs(12345, s)
    da, db, dx = dev(a), dev(b), dev(x)
    c = nttm.mul_add(da, db, dx)
    say("check on demand")
    assert nttm.check_product(c, da, db, dx)
    assert not nttm.check_product(c, da, db)                       # x left out
    for pos in (0, 35000, c.numel() - 2):
        bad = c.clone()
        bad[pos] += 1
        assert not nttm.check_product(bad, da, db, dx)
    assert not nttm.check_product(c[:-1].clone(), da, db, dx) or int(c[-1].item()) == 0
    c0 = nttm.mul(da, db)
    assert nttm.check_product(c0, da, db)


def test_check_residues_agree_with_oracle(nttm):
    """the device residue of a limb array (chunked at 2^16 limbs, weights B^(65536 j)) equals the oracle's: a product
    is accepted exactly when the host residues say c = a b + x mod Q"""
    rng = np.random.default_rng(5)
    for n in (1, 65535, 65536, 65537, 200003):
        a, s = O.gen_limbs(n, 40 + n); b, _ = O.gen_limbs(3, s)
        c = rng.integers(0, O.B, n + 3, dtype=np.uint64)
        want = (O.limbs_mod(a, Q) * O.limbs_mod(b, Q)) % Q == O.limbs_mod(c, Q)
        assert nttm.check_product(dev(c), dev(a), dev(b)) == want == False
        cc = host(nttm.mul(dev(a), dev(b)))
        assert (O.limbs_mod(a, Q) * O.limbs_mod(b, Q)) % Q == O.limbs_mod(cc, Q)
        assert nttm.check_product(dev(cc), dev(a), dev(b))

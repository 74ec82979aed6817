"""ntt_amd: Python binding of libnttm.so, the MI355X multi-prime NTT product of base-10^18 limb integers.

The product itself is C/HIP behind the C ABI of ``include/nttm.h``; this module only loads the shared library and
mirrors its entry points for tests and benchmarks, with device memory held in torch tensors (int64 storage of the
uint64 limbs: every limb is below 10^18 < 2^63).  The names follow the reference's product interface
(``rns_mul_batch`` over ``rns_prod`` records, the one-device product, the ``ntt_*`` primitives; E/rns_mul.h:50-56,
E/ntt.h:37-66 of cecoppola/ntt).  Nothing here computes on the CPU: when the library is missing the calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.environ.get("NTTM_LIB") or os.path.join(HERE, "libnttm.so")   # NTTM_LIB: another build of the library
HEADER = os.path.join(ROOT, "include", "nttm.h")

B = 10 ** 18
Y_FULL, Y_BCAST, Y_PAIR = 0, 1, 2
ERRORS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "EBOUND", -5: "ESTATE", -6: "ECHECK", -7: "EIO"}


class NttmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nttm error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Prod(C.Structure):
    """nttm_prod: c = a b (+ x); ncn receives the normalized length of c."""
    _fields_ = [("a", C.c_void_p), ("na", C.c_size_t), ("b", C.c_void_p), ("nb", C.c_size_t),
                ("c", C.c_void_p), ("x", C.c_void_p), ("nx", C.c_size_t), ("ncn", C.c_size_t)]


class BsNode(C.Structure):
    """nttm_bs_node: P and Q of a binary-splitting node (device limbs)."""
    _fields_ = [("p", C.c_void_p), ("pn", C.c_size_t), ("q", C.c_void_p), ("qn", C.c_size_t)]


class KStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_uint64), ("total_ms", C.c_double), ("alg_bytes", C.c_double)]


R3_AUTO, R3_NEVER, R3_FIT = 0, 1, 2


class Config(C.Structure):
    _fields_ = [("nprimes", C.c_int), ("workspace_bytes", C.c_size_t), ("tile_log", C.c_int), ("pts", C.c_int),
                ("radix3", C.c_int), ("check", C.c_int)]


_lib = None


def lib():
    """The loaded libnttm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnttm.so is not built (run __graft_entry__.build() or python -m ntt_amd.build)")
        try:                     # torch's HIP runtime first: the library then binds to that one copy of it
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, sz, u64, i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int
        L.nttm_init.argtypes = [i32, C.POINTER(Config)]; L.nttm_init.restype = i32
        L.nttm_shutdown.argtypes = []; L.nttm_shutdown.restype = None
        L.nttm_last_error.argtypes = []; L.nttm_last_error.restype = C.c_char_p
        L.nttm_prime.argtypes = [i32]; L.nttm_prime.restype = u64
        L.nttm_root.argtypes = [i32, i32]; L.nttm_root.restype = u64
        L.nttm_max_terms.argtypes = [i32]; L.nttm_max_terms.restype = u64
        L.nttm_mul.argtypes = [vp, vp, sz, vp, sz, vp]; L.nttm_mul.restype = i32
        L.nttm_mul_add.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp]; L.nttm_mul_add.restype = i32
        L.nttm_mul_noalloc.argtypes = [vp, vp, sz, vp, sz, vp]; L.nttm_mul_noalloc.restype = i32
        L.nttm_mul_batch.argtypes = [C.POINTER(Prod), sz, vp]; L.nttm_mul_batch.restype = i32
        L.nttm_mul_cut.argtypes = [vp, vp, sz, vp, sz, sz, sz, vp]; L.nttm_mul_cut.restype = i32
        L.nttm_bs_tree.argtypes = [C.POINTER(BsNode), sz, vp, sz, C.POINTER(C.c_size_t), vp, sz,
                                   C.POINTER(C.c_size_t), vp]
        L.nttm_bs_tree.restype = i32
        L.nttm_e_fraction.argtypes = [C.POINTER(C.c_uint64), sz, vp]; L.nttm_e_fraction.restype = i32
        L.nttm_e_last_phase_ms.argtypes = [C.POINTER(C.c_double)] * 4; L.nttm_e_last_phase_ms.restype = i32
        L.nttm_pack60_bytes.argtypes = [sz]; L.nttm_pack60_bytes.restype = u64
        L.nttm_pack60.argtypes = [vp, vp, sz, sz, vp]; L.nttm_pack60.restype = i32
        L.nttm_limbs_mod.argtypes = [C.POINTER(C.c_uint64), vp, sz, C.POINTER(C.c_uint64), i32, vp]
        L.nttm_limbs_mod.restype = i32
        L.nttm_e_write_packed.argtypes = [C.c_char_p, C.POINTER(C.c_uint64), sz, u64, vp]
        L.nttm_e_write_packed.restype = i32
        L.nttm_check_product.argtypes = [vp, sz, vp, sz, vp, sz, vp, sz, vp]; L.nttm_check_product.restype = i32
        L.nttm_mul_each.argtypes = [C.POINTER(Prod), sz, vp]; L.nttm_mul_each.restype = i32
        L.nttm_mul_each_async.argtypes = [C.POINTER(Prod), sz, vp]; L.nttm_mul_each_async.restype = i32
        L.nttm_batch_plan_create.argtypes = [C.POINTER(Prod), sz, i32, C.POINTER(vp), vp]
        L.nttm_batch_plan_create.restype = i32
        L.nttm_batch_plan_run.argtypes = [vp, vp]; L.nttm_batch_plan_run.restype = i32
        L.nttm_batch_plan_ncn.argtypes = [vp, C.POINTER(C.c_size_t), vp]; L.nttm_batch_plan_ncn.restype = i32
        L.nttm_batch_plan_info.argtypes = [vp, C.c_char_p, sz]; L.nttm_batch_plan_info.restype = i32
        L.nttm_batch_plan_destroy.argtypes = [vp]; L.nttm_batch_plan_destroy.restype = None
        L.nttm_ntt_fwd.argtypes = [i32, vp, i32, sz, vp]; L.nttm_ntt_fwd.restype = i32
        L.nttm_ntt_fwd_mfma.argtypes = [i32, vp, i32, sz, vp]; L.nttm_ntt_fwd_mfma.restype = i32
        L.nttm_ntt_inv.argtypes = [i32, vp, i32, sz, vp]; L.nttm_ntt_inv.restype = i32
        L.nttm_pw.argtypes = [i32, vp, vp, sz, vp]; L.nttm_pw.restype = i32
        L.nttm_ntt_inv_pw.argtypes = [i32, vp, vp, i32, i32, sz, vp]; L.nttm_ntt_inv_pw.restype = i32
        L.nttm_ntt_fwd3.argtypes = [i32, vp, i32, sz, vp]; L.nttm_ntt_fwd3.restype = i32
        L.nttm_ntt_inv3.argtypes = [i32, vp, i32, sz, vp]; L.nttm_ntt_inv3.restype = i32
        L.nttm_ntt_inv3_pw.argtypes = [i32, vp, vp, i32, i32, sz, vp]; L.nttm_ntt_inv3_pw.restype = i32
        L.nttm_root3.argtypes = [i32, i32]; L.nttm_root3.restype = u64
        L.nttm_load.argtypes = [i32, vp, vp, sz, sz, vp]; L.nttm_load.restype = i32
        L.nttm_crt.argtypes = [vp, sz, C.POINTER(vp), i32, sz, vp, sz, vp]; L.nttm_crt.restype = i32
        L.nttm_last_phase_ms.argtypes = [C.POINTER(C.c_double)] * 3; L.nttm_last_phase_ms.restype = i32
        L.nttm_profile.argtypes = [i32]; L.nttm_profile.restype = i32
        L.nttm_profile_read.argtypes = [C.POINTER(KStat), i32]; L.nttm_profile_read.restype = i32
        L.nttm_debug_plan.argtypes = [i32, u64, i32, i32, C.POINTER(C.c_int), i32]; L.nttm_debug_plan.restype = i32
        L.nttm_debug_product_plan.argtypes = [i32, u64, i32, C.POINTER(C.c_int), i32]
        L.nttm_debug_product_plan.restype = i32
        L.nttm_debug_product_len.argtypes = [sz, i32, i32, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.nttm_debug_product_len.restype = i32
        L.nttm_debug_addsub.argtypes = [vp, sz, vp, sz, sz, vp, sz, sz, i32, C.POINTER(C.c_int), vp]
        L.nttm_debug_addsub.restype = i32
        L.nttm_debug_gen_limbs.argtypes = [vp, sz, u64, vp]; L.nttm_debug_gen_limbs.restype = i32
        L.nttm_debug_tile_indices.argtypes = [i32, i32, i32, i32, u64, C.POINTER(C.c_uint64)]
        L.nttm_debug_tile_indices.restype = i32
        L.nttm_debug_top_nonzero.argtypes = [vp, sz, vp, vp]; L.nttm_debug_top_nonzero.restype = i32
        L.nttm_debug_mul_cyclic.argtypes = [vp, vp, sz, vp, sz, sz, vp]; L.nttm_debug_mul_cyclic.restype = i32
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise NttmError(rc, lib().nttm_last_error().decode())


def _stream(stream):
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _dev_limbs(t, name, need=0):
    """check that t is a contiguous int64 CUDA tensor of at least `need` limbs (ValueError before any library call)"""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s: a torch tensor is required, got %s" % (name, type(t).__name__))
    if not t.is_cuda:
        raise ValueError("%s: a CUDA tensor is required (device memory), got one on %s" % (name, t.device))
    if t.dtype != torch.int64:
        raise ValueError("%s: int64 limbs are required, got %s" % (name, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s: a contiguous tensor is required" % name)
    if t.numel() < need:
        raise ValueError("%s: %d limbs, at least %d needed" % (name, t.numel(), need))
    return t


def _any_limbs(t, name, need=0):
    """(address, limbs) of a limb array for the products, which take device or host memory: an int64 torch tensor (CUDA
    or CPU) or an int64 / uint64 numpy array; contiguous, at least `need` limbs (ValueError otherwise)"""
    import numpy as np
    if isinstance(t, np.ndarray):
        if t.dtype not in (np.uint64, np.int64):
            raise ValueError("%s: uint64 or int64 limbs are required, got %s" % (name, t.dtype))
        if not t.flags.c_contiguous:
            raise ValueError("%s: a contiguous array is required" % name)
        n = t.size
        addr = t.ctypes.data
    else:
        import torch
        if not isinstance(t, torch.Tensor):
            raise ValueError("%s: a torch tensor or numpy array is required, got %s" % (name, type(t).__name__))
        if t.dtype != torch.int64:
            raise ValueError("%s: int64 limbs are required, got %s" % (name, t.dtype))
        if not t.is_contiguous():
            raise ValueError("%s: a contiguous tensor is required" % name)
        n = t.numel()
        addr = t.data_ptr()
    if n < need:
        raise ValueError("%s: %d limbs, at least %d needed" % (name, n, need))
    return (addr if n else None), n


# ---- host-side queries (no device needed) ---------------------------------------------------------------------------
def prime(i):
    return int(lib().nttm_prime(i))


def root(i, logn):
    return int(lib().nttm_root(i, logn))


def root3(i, logk):
    return int(lib().nttm_root3(i, logk))


def max_terms(np_):
    return int(lib().nttm_max_terms(np_))


def debug_plan(logn, batch=1, last_log=0, tile_log=13):
    out = (C.c_int * 24)()
    n = lib().nttm_debug_plan(logn, batch, tile_log, last_log, out, 24)
    return [tuple(out[4 * i:4 * i + 4]) for i in range(n)]


def debug_product_plan(logn, batch=1, tile_log=0):
    """the (lo, m, logw, logh) passes of a product transform of 2^logn points"""
    out = (C.c_int * 24)()
    n = lib().nttm_debug_product_plan(logn, batch, tile_log, out, 24)
    if n < 0:
        _check(n)
    return [tuple(out[4 * i:4 * i + 4]) for i in range(n)]


def debug_product_len(ncoef, radix3=R3_AUTO, tile_log=0):
    """(logn, r3): the transform length a product of ncoef coefficients takes, 2^logn or 3 2^logn points."""
    logn, r3 = C.c_int(), C.c_int()
    _check(lib().nttm_debug_product_len(ncoef, radix3, tile_log, C.byref(logn), C.byref(r3)))
    return logn.value, r3.value


def debug_tile_indices(lo, m, logw, logh, tile):
    buf = (C.c_uint64 * (1 << (m + logw + logh)))()
    n = lib().nttm_debug_tile_indices(lo, m, logw, logh, tile, buf)
    return list(buf[:n])


# ---- device entry points ------------------------------------------------------------------------------------------------
def init(device=0, nprimes=0, workspace_bytes=0, tile_log=0, pts=0, radix3=R3_AUTO, check=0):
    cfg = Config(nprimes, workspace_bytes, tile_log, pts, radix3, check)
    _check(lib().nttm_init(device, C.byref(cfg)))


def shutdown():
    lib().nttm_shutdown()


def mul(a, b, out=None, stream=None):
    """c = a b for int64 device tensors of limbs; returns c (na + nb limbs)."""
    import torch
    _dev_limbs(a, "a"); _dev_limbs(b, "b")
    na, nb = a.numel(), b.numel()
    c = _dev_limbs(out, "out", na + nb) if out is not None else torch.empty(max(na + nb, 1), dtype=torch.int64, device=a.device)
    _check(lib().nttm_mul(_ptr(c), _ptr(a), na, _ptr(b), nb, _stream(stream)))
    return c[:na + nb]


def mul_noalloc(a, b, out, stream=None):
    """c = a b into out (na + nb limbs, device) through nttm_mul_noalloc: no allocation, products of at most 2^20
    transform points"""
    _dev_limbs(a, "a"); _dev_limbs(b, "b")
    na, nb = a.numel(), b.numel()
    _dev_limbs(out, "out", na + nb)
    _check(lib().nttm_mul_noalloc(_ptr(out), _ptr(a) if na else None, na, _ptr(b) if nb else None, nb, _stream(stream)))
    return out[:na + nb]


def mul_add(a, b, x, out=None, stream=None):
    """c = a b + x (na + nb + 1 limbs, the top one the final carry)."""
    import torch
    _dev_limbs(a, "a"); _dev_limbs(b, "b"); _dev_limbs(x, "x")
    na, nb = a.numel(), b.numel()
    if x.numel() > na + nb:
        raise ValueError("x: %d limbs, at most na + nb = %d" % (x.numel(), na + nb))
    c = _dev_limbs(out, "out", na + nb + 1) if out is not None else torch.empty(na + nb + 1, dtype=torch.int64, device=a.device)
    _check(lib().nttm_mul_add(_ptr(c), _ptr(a), na, _ptr(b), nb, _ptr(x), x.numel(), _stream(stream)))
    return c[:na + nb + 1]


WHOLE = 2 ** 64 - 1          # w of a cut product that keeps every limb (SIZE_MAX)


def mul_cut(a, b, lowcut=0, w=WHOLE, out=None, stream=None):
    """rns_mul_low_db / high_db / band_db analog: the low min(na + nb, w) limbs of a b (operands cut to w limbs)."""
    import torch
    _dev_limbs(a, "a"); _dev_limbs(b, "b")
    na, nb = a.numel(), b.numel()
    n = min(na + nb, w)
    c = _dev_limbs(out, "out", n) if out is not None else torch.empty(max(n, 1), dtype=torch.int64, device=a.device)
    _check(lib().nttm_mul_cut(_ptr(c), _ptr(a) if na else None, na, _ptr(b) if nb else None, nb, lowcut, w,
                              _stream(stream)))
    return c[:n]


def bs_tree(leaves, stream=None, caps=None):
    """binary-splitting product tree over leaves [(P_i, Q_i)] of int64 device tensors: the root (P, Q), normalized.
    caps: (p_cap, q_cap) limbs of the root's buffers (default: room for the root's unnormalized lengths)."""
    import torch
    arr = (BsNode * len(leaves))()
    for i, (p, q) in enumerate(leaves):
        arr[i] = BsNode(p.data_ptr() if p.numel() else None, p.numel(), q.data_ptr() if q.numel() else None, q.numel())
    dev = leaves[0][0].device if leaves[0][0].numel() else leaves[0][1].device
    pcap = sum(p.numel() for p, _ in leaves) + sum(q.numel() for _, q in leaves) + len(leaves)
    qcap = sum(q.numel() for _, q in leaves)
    if caps is not None:
        pcap, qcap = caps
    P = torch.empty(max(pcap, 1), dtype=torch.int64, device=dev)
    Q = torch.empty(max(qcap, 1), dtype=torch.int64, device=dev)
    pn, qn = C.c_size_t(), C.c_size_t()
    _check(lib().nttm_bs_tree(arr, len(leaves), _ptr(P), pcap, C.byref(pn), _ptr(Q), qcap, C.byref(qn), _stream(stream)))
    return P[:pn.value], Q[:qn.value]


def top_nonzero(x, stream=None):
    """normalized length of the int64 device tensor x (the device scan behind the e pipeline's normalization)"""
    import torch
    out = torch.full((1,), -1, dtype=torch.int64, device=x.device)
    _check(lib().nttm_debug_top_nonzero(_ptr(x) if x.numel() else None, x.numel(), _ptr(out), _stream(stream)))
    return int(out.item())


def mul_cyclic(a, b, L, stream=None):
    """the cyclic product of int64 device tensors a, b (at most L limbs each): L + 3 limbs of sum_k c_k B^k, c the
    length-L cyclic convolution of their limbs"""
    import torch
    v = torch.empty(L + 3, dtype=torch.int64, device=a.device if a.numel() else b.device)
    _check(lib().nttm_debug_mul_cyclic(_ptr(v), _ptr(a) if a.numel() else None, a.numel(),
                                       _ptr(b) if b.numel() else None, b.numel(), L, _stream(stream)))
    return v


def e_fraction(m, stream=None):
    """floor((e - 2) 10^(18 m)) as m limbs (numpy uint64, least significant first)."""
    import numpy as np
    out = np.zeros(m, dtype=np.uint64)
    _check(lib().nttm_e_fraction(out.ctypes.data_as(C.POINTER(C.c_uint64)), m, _stream(stream)))
    return out


def e_last_phase_ms():
    """wall ms of the last e_fraction's phases: leaf folding, product tree, reciprocal, quotient"""
    v = [C.c_double() for _ in range(4)]
    _check(lib().nttm_e_last_phase_ms(*[C.byref(x) for x in v]))
    return dict(zip(("seed", "tree", "recip", "quot"), (x.value for x in v)))


def debug_addsub(a, b, nout, sa=0, sb=0, sub=False, stream=None):
    """(r, carry out): r = a B^sa +- b B^sb mod B^nout through the device addition (int64 device tensors)"""
    import torch
    r = torch.empty(max(nout, 1), dtype=torch.int64, device=a.device)
    ge = C.c_int()
    _check(lib().nttm_debug_addsub(_ptr(r), nout, _ptr(a) if a.numel() else None, a.numel(), sa,
                                   _ptr(b) if b.numel() else None, b.numel(), sb, 1 if sub else 0, C.byref(ge),
                                   _stream(stream)))
    return r[:nout], bool(ge.value)


def gen_limbs_dev(n, seed, stream=None):
    """n synthetic limbs generated on the device (an int64 tensor): the splitmix64 stream seeded `seed`, mod 10^18,
    top limb nonzero -- the same limbs as the oracle's and the bench's host generator"""
    import torch
    x = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
    _check(lib().nttm_debug_gen_limbs(_ptr(x), n, seed, _stream(stream)))
    return x[:n]


def limbs_to_ascii(limbs):
    """base-10^18 limbs (least significant first) -> bytes, 18 ASCII digits per limb from the top limb"""
    import numpy as np
    d = np.ascontiguousarray(np.asarray(limbs, dtype=np.uint64)[::-1])
    out = np.empty((len(d), 18), dtype=np.uint8)
    for part, col in (((d // 10 ** 9).astype(np.uint32), 0), ((d % 10 ** 9).astype(np.uint32), 9)):
        for i in range(8, -1, -1):
            out[:, col + i] = (part % 10).astype(np.uint8) + 48
            part //= 10
    return out.tobytes()


T1_PRIMES = (4611686018427388039, 4611686018427388073, 4611686018427388081, 4611686018427388091,
             4611686018427388093, 4611686018427388097, 4611686018427388157, 4611686018427388181)


def pack60(x, lo=0, hi=None, stream=None):
    """limbs [lo, hi) of the int64 device tensor x, most significant first, in the 60-bit packed form: a uint8 device
    tensor of pack60_bytes(hi - lo) bytes"""
    import torch
    hi = x.numel() if hi is None else hi
    nb = int(lib().nttm_pack60_bytes(hi - lo))
    out = torch.empty((nb + 3) // 4 * 4 or 4, dtype=torch.uint8, device=x.device)
    _check(lib().nttm_pack60(_ptr(out), _ptr(x), lo, hi, _stream(stream)))
    return out[:nb]


def limbs_mod(x, qs=T1_PRIMES, stream=None):
    """[sum_k x[k] 10^(18 k) mod q for q in qs] for an int64 device tensor x of limbs"""
    import numpy as np
    q = np.ascontiguousarray(qs, dtype=np.uint64)
    res = np.zeros(len(q), dtype=np.uint64)
    _check(lib().nttm_limbs_mod(res.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(x) if x.numel() else None, x.numel(),
                                q.ctypes.data_as(C.POINTER(C.c_uint64)), len(q), _stream(stream)))
    return [int(v) for v in res]


def e_write_packed(path, frac, d_out, stream=None):
    """the packed digit file of e from e_fraction's limbs (numpy uint64), d_out digits in its text form"""
    import numpy as np
    f = np.ascontiguousarray(frac, dtype=np.uint64)
    _check(lib().nttm_e_write_packed(os.fsencode(path), f.ctypes.data_as(C.POINTER(C.c_uint64)), len(f), d_out,
                                     _stream(stream)))


def e_digits(m, stream=None):
    """e's first 18 m decimal digits after the point, as a string."""
    return limbs_to_ascii(e_fraction(m, stream)).decode()


def make_prods(items):
    """items: list of (a, b, x_or_None, c) -> ctypes array of Prod.  Each array is an int64 torch tensor (CUDA or CPU)
    or an int64 / uint64 numpy array (the products stage host memory through HBM); c needs na + nb limbs, + 1 with x.
    The array keeps the items alive (``arr._keep``): the records hold their addresses."""
    arr = (Prod * len(items))()
    for i, (a, b, x, c) in enumerate(items):
        pa, na = _any_limbs(a, "product %d: a" % i)
        pb, nb = _any_limbs(b, "product %d: b" % i)
        px, nx = (None, 0) if x is None else _any_limbs(x, "product %d: x" % i)
        if x is not None and nx > na + nb:
            raise ValueError("product %d: x has %d limbs, at most na + nb = %d" % (i, nx, na + nb))
        pc, _ = _any_limbs(c, "product %d: c" % i, na + nb + (0 if x is None else 1))
        if x is not None and px is None:
            px = pc                      # an empty x still marks the carry limb: any valid address, never read
        arr[i] = Prod(pa, na, pb, nb, pc, px, nx, 0)
    arr._keep = list(items)
    return arr


def check_product(c, a, b, x=None, stream=None):
    """True when c = a b + x modulo 4611686018427388039 (device tensors; the residue post-condition), else False"""
    rc = lib().nttm_check_product(_ptr(c) if c.numel() else None, c.numel(), _ptr(a) if a.numel() else None, a.numel(),
                                  _ptr(b) if b.numel() else None, b.numel(),
                                  _ptr(x) if x is not None and x.numel() else None, 0 if x is None else x.numel(),
                                  _stream(stream))
    if rc == -6:
        return False
    _check(rc)
    return True


def mul_batch(prods, stream=None):
    """rns_mul_batch analog over a ctypes Prod array (see make_prods); fills ncn."""
    _check(lib().nttm_mul_batch(prods, len(prods), _stream(stream)))
    return prods


def mul_each(prods, stream=None):
    """one product call each (synchronizing per product)"""
    _check(lib().nttm_mul_each(prods, len(prods), _stream(stream)))
    return prods


def mul_each_async(prods, stream=None):
    """one product call each, launched back to back, one synchronization at the end (device memory only)"""
    _check(lib().nttm_mul_each_async(prods, len(prods), _stream(stream)))
    return prods


class BatchPlan:
    """a batch of products (a ctypes Prod array, see make_prods) planned once -- schedule autotuned, descriptors
    uploaded, launches captured as one hipGraph -- and run many times"""

    def __init__(self, prods, autotune=True, stream=None):
        """autotune: False (the library's defaults), True (every candidate timed; the choices are remembered) or
        "reuse" (a choice remembered for a group of the same length and size class is taken untimed)"""
        self.prods = prods                                       # the records (their graph has the addresses built in)
        self._keep = getattr(prods, "_keep", None)               # and the arrays they point to, kept alive with the plan
        self.n = len(prods)
        h = C.c_void_p()
        mode = 2 if autotune == "reuse" else (1 if autotune else 0)
        _check(lib().nttm_batch_plan_create(prods, self.n, mode, C.byref(h), _stream(stream)))
        self.h = h

    def run(self, stream=None):
        _check(lib().nttm_batch_plan_run(self.h, _stream(stream)))

    def ncn(self, stream=None):
        out = (C.c_size_t * max(self.n, 1))()
        _check(lib().nttm_batch_plan_ncn(self.h, out, _stream(stream)))
        return [int(out[i]) for i in range(self.n)]

    def info(self):
        buf = C.create_string_buffer(1 << 16)
        _check(lib().nttm_batch_plan_info(self.h, buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if self.h:
            lib().nttm_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt_fwd(prime_idx, x, logn, batch=1, stream=None):
    _check(lib().nttm_ntt_fwd(prime_idx, _ptr(x), logn, batch, _stream(stream)))


def ntt_fwd_mfma(prime_idx, x, logn, batch=1, stream=None):
    """nttm_ntt_fwd with its first four levels on INT8 matrix cores (identical output)"""
    _check(lib().nttm_ntt_fwd_mfma(prime_idx, _ptr(x), logn, batch, _stream(stream)))


def ntt_inv(prime_idx, x, logn, batch=1, stream=None):
    _check(lib().nttm_ntt_inv(prime_idx, _ptr(x), logn, batch, _stream(stream)))


def ntt_inv_pw(prime_idx, x, y, ymode, logn, batch=1, stream=None):
    _check(lib().nttm_ntt_inv_pw(prime_idx, _ptr(x), _ptr(y), ymode, logn, batch, _stream(stream)))


def ntt_fwd3(prime_idx, x, logk, batch=1, stream=None):
    """transforms of 3 2^logk points: third r, position brev(t) receives X[3t + r] (ntt_fwd3)."""
    _check(lib().nttm_ntt_fwd3(prime_idx, _ptr(x), logk, batch, _stream(stream)))


def ntt_inv3(prime_idx, x, logk, batch=1, stream=None):
    _check(lib().nttm_ntt_inv3(prime_idx, _ptr(x), logk, batch, _stream(stream)))


def ntt_inv3_pw(prime_idx, x, y, ymode, logk, batch=1, stream=None):
    _check(lib().nttm_ntt_inv3_pw(prime_idx, _ptr(x), _ptr(y), ymode, logk, batch, _stream(stream)))


def pw(prime_idx, x, y, stream=None):
    _check(lib().nttm_pw(prime_idx, _ptr(x), _ptr(y), x.numel(), _stream(stream)))


def load(prime_idx, dst, src, stream=None):
    _check(lib().nttm_load(prime_idx, _ptr(dst), _ptr(src) if src.numel() else None, src.numel(), dst.numel(),
                           _stream(stream)))


def crt(planes, ncoef, nc, x=None, stream=None):
    """limbs of sum_k CRT(planes[.][k]) B^k (+ x) for canonical residue planes; returns c (nc limbs)."""
    import torch
    arr = (C.c_void_p * len(planes))(*[p.data_ptr() for p in planes])
    c = torch.empty(max(nc, 1), dtype=torch.int64, device=planes[0].device)
    _check(lib().nttm_crt(_ptr(c), nc, arr, len(planes), ncoef, _ptr(x), 0 if x is None else x.numel(), _stream(stream)))
    return c[:nc]


def last_phase_ms():
    f, i, c = C.c_double(), C.c_double(), C.c_double()
    lib().nttm_last_phase_ms(C.byref(f), C.byref(i), C.byref(c))
    return f.value, i.value, c.value


def profile(enable=True):
    """Bracket every kernel launch with HIP events (clears earlier records)."""
    _check(lib().nttm_profile(1 if enable else 0))


def profile_read():
    """{kernel family: (launches, total_ms, alg_bytes)} of the launches recorded since profile(True)."""
    arr = (KStat * 64)()
    n = lib().nttm_profile_read(arr, 64)
    if n < 0:
        _check(n)
    return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms), float(arr[i].alg_bytes)) for i in range(n)}

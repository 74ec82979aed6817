"""Helpers shared by the GPU tests: moving uint64 limb arrays to and from int64 device tensors, and printing the case
before each GPU step so a failing run names it."""
import sys

import numpy as np


def dev(a):
    import torch
    a = np.ascontiguousarray(a, dtype=np.uint64)
    if a.size == 0:
        return torch.empty(0, dtype=torch.int64, device="cuda")
    return torch.from_numpy(a.view(np.int64).copy()).cuda()


def host(t):
    return t.detach().cpu().numpy().view(np.uint64).copy()


def say(*args):
    print("[case]", *args, flush=True)
    sys.stdout.flush()

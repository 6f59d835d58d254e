"""Helpers shared by the GPU tests: device copies of Fortran-ordered arrays."""
import numpy as np
import torch


def to_dev(a):
    """Column-major numpy array -> device tensor holding the same bytes."""
    a = np.asfortranarray(a)
    t = torch.from_numpy(a.T.copy())  # row-major (n, m) == column-major (m, n)
    return t.cuda()


def from_dev(t, shape):
    """Device tensor written column-major -> numpy (m, n) Fortran array."""
    m, n = shape
    return np.asfortranarray(t.cpu().numpy().reshape(n, m).T)


def ivec(n):
    return torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")

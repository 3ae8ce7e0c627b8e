"""Shared helpers for the test suite."""
import numpy as np


def rel_l1(a, b, mask=None):
    """Relative L1 difference sum|a-b| / sum|b| (over mask when given)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a = a[..., mask]
        b = b[..., mask]
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


def updated_mask(bctype):
    ny, nx = bctype.shape
    m = np.zeros((ny, nx), dtype=bool)
    m[3:ny - 3, 3:nx - 3] = True
    return m & (bctype == 0)

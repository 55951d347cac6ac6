"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

# D3Q19 (descriptor/definition/common.h:166-199)
C = np.array([
    [0, 0, 0],
    [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
    [-1, 0, 1], [0, -1, -1], [0, -1, 1],
    [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, -1, 0], [1, 0, 1],
    [1, 0, -1], [0, 1, 1], [0, 1, -1],
], dtype=np.int64)
T = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)
OPP = np.array([0] + list(range(10, 19)) + list(range(1, 10)))


def rel_err_unshifted(f_test: np.ndarray, f_ref: np.ndarray) -> float:
    """max_i,n |f_test - f_ref| / |f_ref + t_i|: relative error of the
    unshifted populations f_i = f_bar_i + t_i."""
    ref = f_ref.astype(np.float64)
    return float(np.max(np.abs(f_test.astype(np.float64) - ref) / np.abs(ref + T[:, None])))


def core_cells(case) -> np.ndarray:
    return case.core_mask().reshape(-1)


def velocity(f: np.ndarray) -> np.ndarray:
    """u = j / rho from shifted populations (19, n) -> (n, 3)."""
    f = f.astype(np.float64)
    rho = f.sum(axis=0) + 1.0
    j = C.T.astype(np.float64) @ f
    return (j / rho).T


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

"""Helpers shared by the GPU parity tests."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def say(msg: str) -> None:
    print(msg, flush=True)


def rel_l2_scaled(a, b, scale: float) -> float:
    """||a - b|| / max(||b||, scale): relative L2 that stays meaningful for a
    component that is zero up to roundoff (e.g. v in a y-symmetric box), where
    `scale` is the norm of the whole velocity field."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), scale, 1e-300))


def vel_norm(u, v, w) -> float:
    return float(np.sqrt(np.linalg.norm(np.asarray(u, np.float64)) ** 2
                         + np.linalg.norm(np.asarray(v, np.float64)) ** 2
                         + np.linalg.norm(np.asarray(w, np.float64)) ** 2))

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


# A velocity component whose norm is below this fraction of the whole velocity
# field's is zero up to roundoff (v in a y-symmetric box): its own-norm error is
# noise, and only the velocity-scaled figure is meaningful for it.  In FP32 the
# threshold is ROUNDOFF_FP32: a component of 1e-6 of the velocity norm is some
# 16 FP32 ulps of the largest one, and T - baseT that small beside T sits at
# the FP32 resolution of T itself, so their own-norm errors are rounding noise.
ROUNDOFF_COMPONENT = 1e-9
ROUNDOFF_FP32 = 1e-6


def component_errors(got, want, base_t: float, v_T: float,
                     roundoff: float = ROUNDOFF_COMPONENT) -> dict:
    """Per-component relative L2 errors of a state against a reference state.

    got:  (u, v, w, T) of the state checked; want: (u, v, w, T) of the reference.
    Returns {name: (own, scaled)}: `own` is each component's error relative to
    its own signal -- u, v, w to their own norms, T as T - baseT, p = v_T T as
    p - v_T baseT (T stays close to baseT, so relative to T itself an error
    would hide in the constant) -- and `scaled` is the velocity-norm-scaled
    figure for u, v, w (T and p relative to themselves).  `own` is None for a
    component that is zero up to `roundoff` in the reference (a velocity
    component against the velocity norm, T - baseT against T).
    """
    u, v, w, T = (np.asarray(x, np.float64) for x in got)
    ru, rv, rw, rT = (np.asarray(x, np.float64) for x in want)
    vs = vel_norm(ru, rv, rw)
    out = {}
    for nm, a, b in (("u", u, ru), ("v", v, rv), ("w", w, rw)):
        nb = float(np.linalg.norm(b))
        own = float(np.linalg.norm(a - b) / nb) if nb > roundoff * vs and nb > 0 else None
        out[nm] = (own, rel_l2_scaled(a, b, vs))
    dT = rT - base_t
    ndT = float(np.linalg.norm(dT))
    eT = float(np.linalg.norm(T - rT))
    t_ok = ndT > 0 and ndT > roundoff * float(np.linalg.norm(rT))
    out["T"] = (eT / ndT if t_ok else None, rel_l2(T, rT))
    # p = v_T T: the error of p - v_T baseT equals v_T times the error of T - baseT
    p, rp = v_T * T, v_T * rT
    dp = rp - v_T * base_t
    ndp = float(np.linalg.norm(dp))
    out["p"] = (float(np.linalg.norm(p - rp)) / ndp if t_ok else None, rel_l2(p, rp))
    return out


def format_errors(errs: dict) -> str:
    parts = []
    for nm, (own, sc) in errs.items():
        o = "roundoff" if own is None else f"{own:.2e}"
        parts.append(f"{nm} own {o} scaled {sc:.2e}")
    return "; ".join(parts)


def check_fp32_against_spread(gpu: dict, oracle32: dict, factor: float = 2.0,
                              floor: float = 1e-7) -> list:
    """FP32 bar set by the FP32 oracle's own spread against the FP64 oracle:
    each component's own-norm error of the GPU result may be at most `factor`
    times the FP32 oracle's on that component (`floor` for a component the
    FP32 oracle happens to hit almost exactly).  Returns the failures."""
    bad = []
    for nm, (own, _) in gpu.items():
        ref = oracle32[nm][0]
        if own is None or ref is None:
            continue
        bar = max(factor * ref, floor)
        if not own <= bar:
            bad.append((nm, own, bar))
    return bad

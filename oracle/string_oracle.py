"""CPU restatement of one String iteration (test infrastructure only).

Follows the reference solver's kernels (native/string/source_code/global.h):
s0_grad_cal 1281-1324 (force floored to a 1e-15 grid), s0_grad_length_sqrt_cal
1328-1341, s0_new_cal 1347-1443 (separate translational / rotational norms),
s1_fix_modify_upgrade 1456-1647 (applied in image order), s1_frac2cart /
s1_length_prep / cumulative sums / s1_2_s2 1653-1846 (arc-length
reparametrization of a, b, c), check_s2 1854-1896, smooth_angle 1905-2085,
diff_s_prep + check_signal 2091-2163 (legacy_policy.h:9-13), copy2s0 2169-2240.
The force is delta_k * dE/dx_k from the FP64 analytic gradient (oracle
pose_grad) by the chain rule, where the reference takes forward differences.

Pinned by tests/test_string_oracle.py: the energy equals the sum of the
oracle's triclinic pair (itself bit-equal to the reference header) over sites
and atoms, and the gradient matches its central differences; smoothing
against its closed form and the 0 / 2 pi wrap branches; after a step the
endpoints are fixed, the last image is the periodic replica of the first and
the chords are equal; exactly the images whose energy rose are reverted; a
converged string stops only on a check iteration (planted energy functions).
"""
from __future__ import annotations

import numpy as np

from . import vext_oracle as O

ROUND = 1e-15
SMOOTH = 1e-4


def _floor(v):
    return np.floor(v / ROUND) * ROUND


def director(beta, gamma):
    return np.stack([np.cos(gamma) * np.cos(beta), np.sin(gamma) * np.cos(beta), -np.sin(beta)], -1)


def energy_grad(case, x, setup=None):
    """E and dE/d(a, b, c, alpha, beta, gamma) of poses x (n, 6)."""
    cell = np.asarray(case["cell"])
    be, ga = x[:, 4], x[:, 5]
    e, gc, gu = O.pose_grad(case, x[:, :3] @ cell, director(be, ga), setup=setup)
    g = np.zeros_like(x)
    g[:, :3] = gc @ cell.T
    g[:, 4] = (gu * np.stack([-np.cos(ga) * np.sin(be), -np.sin(ga) * np.sin(be), -np.cos(be)], -1)).sum(-1)
    g[:, 5] = (gu * np.stack([-np.sin(ga) * np.cos(be), np.cos(ga) * np.cos(be), np.zeros_like(be)], -1)).sum(-1)
    return e, g


def smooth(ang):
    n = len(ang)
    out = ang.copy()
    for i in range(1, n - 1):
        for k in range(3):
            c, l, r = ang[i, k], ang[i - 1, k], ang[i + 1, k]
            dl, dr = l - c, r - c
            if dl > np.pi:
                w = -2.0 if dr > np.pi else -1.0
            elif dl < -np.pi:
                w = 2.0 if dr < -np.pi else 0.0
            else:
                w = -1.0 if dr > np.pi else (1.0 if dr < -np.pi else 0.0)
            out[i, k] = (1 - SMOOTH) * c + 0.5 * SMOOTH * (l + r + w * 2 * np.pi)
    return out


def iterate(case, s0, direction, it, move_frac=1e-4, move_angle=np.radians(1.0), setup=None):
    """One iteration of one string; returns (new s0, converged)."""
    cell = np.asarray(case["cell"])
    n = len(s0)
    v0, grad = energy_grad(case, s0, setup)
    mv = np.array([move_frac] * 3 + [move_angle] * 3)
    g = _floor(mv * grad)
    nt = _floor(np.sqrt((g[:, :3] ** 2).sum(1)))
    nr = _floor(np.sqrt((g[:, 3:] ** 2).sum(1)))
    s1 = s0.copy()
    for i in range(n):
        for k in range(6):
            nn = nt[i] if k < 3 else nr[i]
            step = g[i, k] if nn < mv[k] else _floor(g[i, k] / nn)
            s1[i, k] = s0[i, k] - mv[k] * step
    d = direction - 1
    s1[0, d] = s0[0, d]
    s1[n - 1, d] = s0[n - 1, d]
    for k in range(6):
        if k != d:
            s1[n - 1, k] = s1[0, k]
    for i in range(1, n - 1):
        src = None
        if s1[i, d] < 0:
            src = next((j for j in range(i + 1, n) if s1[j, d] > 0), None)
        elif s1[i, d] > 1:
            src = next((j for j in range(i - 1, -1, -1) if s1[j, d] < 1), None)
        if src is not None:
            s1[i] = s1[src]
    cart = s1[:, :3] @ cell
    cum = np.concatenate([[0.0], np.cumsum(np.sqrt(((cart[1:] - cart[:-1]) ** 2).sum(1)))])
    total = cum[-1]
    s2 = s1.copy()
    for i in range(n):
        t = i / (n - 1)
        ii = 0
        if total > 0:
            ii = max(j for j in range(n - 1) if cum[j] / total <= t)
        f0 = cum[ii] / total if total > 0 else 0.0
        f1 = cum[ii + 1] / total if total > 0 else 0.0
        for k in range(3):
            a0, a1 = s1[ii, k], s1[ii + 1, k]
            s2[i, k] = a0 + (t - f0) * ((a1 - a0) / (f1 - f0)) if f1 > f0 else a0
    v2, _ = energy_grad(case, s2, setup)
    ang = np.where((v2 > v0)[:, None], s0[:, 3:], s2[:, 3:])
    tot_c = np.sqrt(((s0[:, :3] - s2[:, :3]) ** 2).sum(1)).sum()
    tot_o = np.sqrt(((s0[:, 3:] - ang) ** 2).sum(1)).sum()
    conv_c = 0.30 * move_frac * n * np.sqrt(3)
    conv_o = 0.30 * move_angle * n * np.sqrt(3)
    if it % 200 == 0 and it >= 200 and tot_c < conv_c and tot_o < conv_o:
        return s0.copy(), True
    out = s2.copy()
    sm = smooth(ang)
    sm = np.where(sm > 2 * np.pi, sm - 2 * np.pi, np.where(sm < 0, sm + 2 * np.pi, sm))
    out[:, 3:] = sm
    return out, False

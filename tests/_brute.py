"""Independent NumPy brute force of the NAIVE formulation (naive_kernels.py:7-80), written
as dense pair matrices rather than loops.  It cross-checks the C oracle and gives the
per-particle magnitude sums used to normalise force errors (a pressure or viscosity sum
cancels between neighbours, so an error is measured against sum_j |term_ij|).
"""
from __future__ import annotations

import math

import numpy as np


def naive_terms(pos, vel, c, rho=None, block=512):
    pos = np.asarray(pos, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    n = pos.shape[0]
    h, h2 = c.h, c.h ** 2
    wc = 315.0 / (64.0 * math.pi * h ** 9)
    gc = -45.0 / (math.pi * h ** 6)
    lc = 45.0 / (math.pi * h ** 6)
    if rho is None:
        rho = np.empty(n)
        for s in range(0, n, block):
            d = pos[s:s + block, None, :] - pos[None, :, :]
            r2 = np.sum(d * d, axis=2)
            rho[s:s + block] = c.mass * np.sum(np.where(r2 <= h2, wc * (h2 - r2) ** 3, 0.0), axis=1)
    rho = np.asarray(rho, dtype=np.float64)
    p = c.k * (rho - c.rho0)
    P = np.zeros((n, 3))
    V = np.zeros((n, 3))
    sP = np.zeros(n)
    sV = np.zeros(n)
    for s in range(0, n, block):
        e = min(n, s + block)
        d = pos[s:e, None, :] - pos[None, :, :]
        r2 = np.sum(d * d, axis=2)
        mask = r2 <= h2
        mask[np.arange(e - s), np.arange(s, e)] = False
        r = np.sqrt(np.where(mask, r2, 1.0))
        g = np.where(mask, gc * (h - r) ** 2 / r, 0.0)[:, :, None] * d
        coef = (p[s:e, None] + p[None, :]) / (2.0 * rho[None, :])
        P[s:e] = c.mass * np.sum(coef[:, :, None] * g, axis=1)
        sP[s:e] = c.mass * np.sum(np.abs(coef) * np.linalg.norm(g, axis=2), axis=1)
        lap = np.where(mask, lc * (h - r), 0.0)
        dv = vel[None, :, :] - vel[s:e, None, :]
        wv = (c.mass * c.visc * lap / rho[None, :])[:, :, None]
        V[s:e] = np.sum(wv * dv, axis=1)
        sV[s:e] = np.sum(np.abs(wv[:, :, 0]) * np.linalg.norm(dv, axis=2), axis=1)
    return rho, P, V, sP, sV


def neighbour_counts(pos, h):
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    out = np.zeros(n, dtype=np.int64)
    for s in range(0, n, 512):
        d = pos[s:s + 512, None, :] - pos[None, :, :]
        out[s:s + 512] = np.sum(np.sum(d * d, axis=2) <= h * h, axis=1)
    return out


def near_boundary(pos, h, rel):
    """Particles with a partner at distance within rel*h of h (either side of the cut-off)."""
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    out = np.zeros(n, dtype=bool)
    for s in range(0, n, 512):
        d = np.sqrt(np.sum((pos[s:s + 512, None, :] - pos[None, :, :]) ** 2, axis=2))
        out[s:s + 512] = np.any(np.abs(d - h) < rel * h, axis=1)
    return out

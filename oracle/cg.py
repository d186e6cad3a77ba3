"""Conjugate gradient as IterativeSolvers.jl 0.9 runs it (test infrastructure).

IterativeSolvers.jl is a third-party dependency (Project.toml:7,17, compat
0.9) that is not vendored in the reference; its call sites are
test/reconstruct_radial_gpu.jl:99 (``cg(A, vec(b); maxiter)``) and
src/CoilMaps.jl:136 (``cg!(..., reltol=0)``).  Restated recurrence, with x0 = 0
(``cg`` passes ``initially_zero = true``):

    r = b; u = 0; residual = ||b||; prev = 1; tol = max(reltol*residual, abstol)
    repeat while not (it >= maxiter or residual <= tol):
        beta = residual^2 / prev^2;  u = r + beta u
        c = A u;  alpha = residual^2 / dot(u, c)     (dot conjugates u)
        x += alpha u;  r -= alpha c
        prev = residual;  residual = ||r||
"""
from __future__ import annotations

import math

import numpy as np

REL_TOL_F32 = math.sqrt(float(np.finfo(np.float32).eps))


def cg(apply, b, maxiter, reltol=REL_TOL_F32, abstol=0.0, x0=None, history=None):
    b = np.asarray(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, copy=True)
    if x0 is None:
        r = b.copy()
        residual = float(np.linalg.norm(r))
    else:
        r = b - apply(x)
        residual = float(np.linalg.norm(r))
    tol = max(reltol * residual, abstol)
    u = np.zeros_like(b)
    prev = 1.0
    it = 0
    while not (it >= maxiter or residual <= tol):
        beta = residual ** 2 / prev ** 2
        u = r + beta * u
        c = apply(u)
        alpha = residual ** 2 / np.vdot(u.reshape(-1), c.reshape(-1))
        x = x + alpha * u
        r = r - alpha * c
        prev = residual
        residual = float(np.linalg.norm(r))
        it += 1
        if history is not None:
            history.append(residual)
    return x

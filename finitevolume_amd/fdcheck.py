"""Finite-difference estimates of a directional derivative, for checking
exact derivatives (the adjoint, the forward-mode derivative).

g(s) is the change of the objective along the direction at step s, with
g(0) = 0 by construction (the caller sums lambda * (R(s) - R(0)) entry by
entry, so entries the step leaves unchanged cancel exactly instead of
adding the rounding of the full sum).  For each step of a sweep the
estimates are the central difference, its Richardson extrapolation (error
O(h^4)) and the two second-order one-sided differences; the one-sided ones
are there because the limiter's min/max and the HLLC branch make the
residual piecewise smooth: on a tie (a wall ghost equal to its cell in one
component, for instance) the exact derivative is one of the two one-sided
slopes and the central difference is their mean.  The result is the best
agreement with the exact value over all estimates and steps."""
from __future__ import annotations

from typing import Callable, Iterable, Tuple


def best_fd(g: Callable[[float], float], exact: float, steps: Iterable[float]) -> Tuple[float, float, str]:
    """(best |estimate - exact|, the step it was reached at, which estimate)."""
    return fd_sweep(g, exact, steps)[:3]


def fd_sweep(g: Callable[[float], float], exact: float, steps: Iterable[float]) -> Tuple[float, float, str, bool]:
    """best_fd's result plus whether the point is a kink: the two one-sided
    slopes disagree by more than 1e-6 of the larger at every step of the
    sweep (a tie in the limiter's min/max or in the HLLC wave choice, where
    the exact derivative follows the tie-breaking rule and may equal neither
    slope)."""
    best = (float("inf"), 0.0, "")
    gap = float("inf")
    for h in steps:
        gp, gm, gp2, gm2 = g(h), g(-h), g(2 * h), g(-2 * h)
        c1 = (gp - gm) / (2 * h)
        c2 = (gp2 - gm2) / (4 * h)
        est = {"central": c1, "richardson": (4 * c1 - c2) / 3, "forward": (4 * gp - gp2) / (2 * h),
               "backward": (gm2 - 4 * gm) / (2 * h)}
        for k, v in est.items():
            e = abs(v - exact)
            if e < best[0]:
                best = (e, h, k)
        f, b = est["forward"], est["backward"]
        gap = min(gap, abs(f - b) / max(abs(f), abs(b), 1e-300))
    return best + (gap > 1e-6,)

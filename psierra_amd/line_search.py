"""Backtracking line search with pSierra's acceptance rules (host logic).

Same rules and constants as ContactAwareLineSearch
(src/psierra/analysis/line_search.py:68-260) -- accept on direct
convergence, on any decrease of ||R||, on the Armijo condition
||R_a||^2 <= (1 - 2 c1 a) ||R_0||^2, or (with contact) on a bounded full
step; otherwise halve, at most 16 times, then fall back to the best trial
-- and the same exception classifier (line_search.py:10-33), so inversion
errors raised by the GPU wrapper backtrack exactly as pSierra's do.
Each trial is one force-only GPU assembly plus the FP64 residual norm.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

_INVERSION_KEYWORDS = (
    "element inversion",
    "negative jacobian",
    "det(f)",
    "outside admissible",
    "admissible range",
    "zero surface area",
    "degenerate",
)


def is_element_inversion_error(exc: BaseException) -> bool:
    if isinstance(exc, FloatingPointError):
        return True
    return isinstance(exc, ValueError) and any(k in str(exc).lower() for k in _INVERSION_KEYWORDS)


@dataclass(frozen=True)
class LineSearchResult:
    success: bool
    alpha: float
    candidate_norm: float
    iterations: int
    element_inversion_encountered: bool
    accepted_as_contact_fallback: bool
    rejection_reason: str | None = None


class ContactAwareLineSearch:
    def __init__(self, max_iterations=16, contraction_factor=0.5, armijo_c1=2.0e-4,
                 full_step_tolerance=1.5, contact_fallback_tolerance=5.0):
        if max_iterations < 1:
            raise ValueError("max_iterations must be at least 1")
        if not 0.0 < contraction_factor < 1.0:
            raise ValueError("contraction_factor must be in (0, 1)")
        if not 0.0 < armijo_c1 < 0.5:
            raise ValueError("armijo_c1 must be in (0, 0.5)")
        if full_step_tolerance < 1.0:
            raise ValueError("full_step_tolerance must be at least 1.0")
        if contact_fallback_tolerance < 1.0:
            raise ValueError("contact_fallback_tolerance must be at least 1.0")
        self.max_iterations = int(max_iterations)
        self.contraction_factor = float(contraction_factor)
        self.armijo_c1 = float(armijo_c1)
        self.full_step_tolerance = float(full_step_tolerance)
        self.contact_fallback_tolerance = float(contact_fallback_tolerance)

    def search(self, eval_fn, current_norm, inner_tolerance, *, contact_active=False, reference_norm=None):
        scale = current_norm if reference_norm is None else max(current_norm, reference_norm)
        alpha = 1.0
        best_norm, best_alpha = math.inf, 1.0
        inverted_any = False
        for it in range(1, self.max_iterations + 1):
            try:
                cand = float(eval_fn(alpha))
            except Exception as exc:  # classified like the reference, others propagate
                if not is_element_inversion_error(exc):
                    raise
                cand = math.inf
                inverted_any = True
            if not math.isfinite(cand):
                inverted_any = True
                alpha *= self.contraction_factor
                continue
            if cand < best_norm:
                best_norm, best_alpha = cand, alpha
            accept = (
                cand <= inner_tolerance
                or cand <= current_norm
                or cand * cand <= (1.0 - 2.0 * self.armijo_c1 * alpha) * current_norm * current_norm
                or (contact_active and it == 1 and alpha == 1.0 and cand < self.full_step_tolerance * scale)
            )
            if accept:
                return LineSearchResult(True, alpha, cand, it, inverted_any, False)
            alpha *= self.contraction_factor
        if math.isfinite(best_norm) and best_norm < self.contact_fallback_tolerance * scale and (
            contact_active or not inverted_any
        ):
            return LineSearchResult(True, best_alpha, best_norm, self.max_iterations, inverted_any, True)
        if not math.isfinite(best_norm) and inverted_any:
            reason = "Element inversion detected across all evaluated step lengths"
        else:
            reason = (
                f"Residual {best_norm:.4e} exceeded fallback tolerance "
                f"{self.contact_fallback_tolerance * scale:.4e}"
            )
        return LineSearchResult(False, best_alpha, best_norm, self.max_iterations, inverted_any, False, reason)

"""The reference's scoring harness, formula version 1 (src/verification/Scoring.jl,
src/invent/Candidate.jl, src/verification/schema_keys.jl): absolute and relative suite scores
from case dictionaries, the order/dissipation trade-off test and the candidate classification
of an invented method against a baseline report.  Case dictionaries use the reference's JSON
keys, so a report written by the reference and one written here score identically."""
from __future__ import annotations

import math
from typing import Any, Dict, Iterable, List, Mapping, Optional, Tuple

SCHEMA_VERSION = 1              # schema_keys.jl:5
SCORING_FORMULA_VERSION = 1     # schema_keys.jl:6
DEFAULT_SCORING_WEIGHTS = {     # schema_keys.jl:7-12
    "order_preservation": 0.30,
    "dissipation": 0.25,
    "shock_quality": 0.25,
    "robustness": 0.20,
}
DEFAULT_SCORE_MARGIN = 0.02     # invent margin delta of classify_candidate

D_REF_ABS = 0.2                 # Scoring.jl:3-6
DELTA_REF_ABS = 8.0             # shock thickness in solution-point spacings
ETA_REF_ABS = 0.1
SCORE_EPS = 1e-30


def clip01(x) -> float:
    return min(max(float(x), 0.0), 1.0)


def _missing(v) -> bool:
    return v is None


def _is(v, want: bool) -> bool:
    """Julia's `=== true` / `=== false`: only a real Bool matches."""
    return isinstance(v, bool) and v is want


def _order_score(cases) -> float:
    order_cases = [c for c in cases if c.get("case_type", "") == "smooth_order"]
    if not order_cases:
        return 1.0
    return 1.0 if all(_is(c.get("order_pass", False), True) for c in order_cases) else 0.0


def _robust_score(cases) -> float:
    fail = any(_is(c.get("diverged", False), True) or _is(c.get("nan_detected", False), True)
               or _is(c.get("positivity_ok", True), False) for c in cases)
    return 0.0 if fail else 1.0


def _composite(S_order, S_diss, S_shock, S_robust) -> float:
    w = DEFAULT_SCORING_WEIGHTS
    return (w["order_preservation"] * S_order + w["dissipation"] * S_diss
            + w["shock_quality"] * S_shock + w["robustness"] * S_robust)


def _mean(v: List[float]) -> float:
    if not v:
        return 1.0
    s = 0.0
    for x in v:      # left-to-right, as Julia's sum over a short vector
        s += x
    return s / len(v)


def _scores(S_order, diss, shock, S_robust) -> Dict[str, float]:
    S_diss, S_shock = _mean(diss), _mean(shock)
    return {"order_preservation": S_order, "dissipation": S_diss, "shock_quality": S_shock,
            "robustness": S_robust, "composite": _composite(S_order, S_diss, S_shock, S_robust)}


def score_suite_absolute(cases: Iterable[Mapping[str, Any]]) -> Dict[str, float]:
    """S_order (all smooth-order cases pass), S_diss and S_shock (means over the discontinuous
    cases of the absolute maps), S_robust and the weighted composite (Scoring.jl:16-73)."""
    cases = list(cases)
    diss, shock = [], []
    for c in cases:
        if c.get("case_type", "") != "discontinuous":
            continue
        dex = c.get("excess_dissipation")
        eta = float(c.get("overshoot", 0.0))
        if _missing(dex):
            diss.append(clip01(1 - eta / ETA_REF_ABS))
        else:
            diss.append(clip01(1 - float(dex) / D_REF_ABS))
        delta = c.get("shock_thickness")
        if _missing(delta) or not math.isfinite(float(delta)):
            s_delta = 1.0
        else:
            s_delta = clip01(1 - float(delta) / DELTA_REF_ABS)
        shock.append(0.5 * s_delta + 0.5 * clip01(1 - eta / ETA_REF_ABS))
    return _scores(_order_score(cases), diss, shock, _robust_score(cases))


def _match_baseline(mc, b_disc):
    """Pair a method case with the baseline case of the same problem (Candidate.jl:23-37)."""
    mname = mc.get("name", "")
    for bc in b_disc:
        bname = bc.get("name", "")
        if mname.startswith("sod") and bname.startswith("sod"):
            return bc
        if mname.startswith("shu_osher") and bname.startswith("shu_osher"):
            return bc
    return b_disc[0] if b_disc else None


def score_suite_relative(method_cases, baseline_cases) -> Dict[str, float]:
    """Relative maps against a baseline's cases (Candidate.jl:8-96)."""
    method_cases, baseline_cases = list(method_cases), list(baseline_cases)
    m_disc = [c for c in method_cases if c.get("case_type", "") == "discontinuous"]
    b_disc = [c for c in baseline_cases if c.get("case_type", "") == "discontinuous"]
    diss, shock = [], []
    for mc in m_disc:
        bc = _match_baseline(mc, b_disc)
        eta_m = float(mc.get("overshoot", 0.0))
        dex_m = mc.get("excess_dissipation")
        delta_m = mc.get("shock_thickness")
        if bc is None:
            diss.append(clip01(1 - eta_m / ETA_REF_ABS) if dex_m is None
                        else clip01(1 - float(dex_m) / D_REF_ABS))
            s_delta = (1.0 if delta_m is None or not math.isfinite(float(delta_m))
                       else clip01(1 - float(delta_m) / DELTA_REF_ABS))
            shock.append(0.5 * s_delta + 0.5 * clip01(1 - eta_m / ETA_REF_ABS))
            continue
        eta_b = float(bc.get("overshoot", 0.0))
        dex_b = bc.get("excess_dissipation")
        delta_b = bc.get("shock_thickness")
        if dex_m is None or dex_b is None:
            diss.append(clip01(eta_b / (eta_m + SCORE_EPS)))
        else:
            diss.append(clip01(float(dex_b) / (float(dex_m) + SCORE_EPS)))
        if (delta_m is None or delta_b is None or not math.isfinite(float(delta_m))
                or not math.isfinite(float(delta_b))):
            s_delta = 1.0
        else:
            s_delta = clip01(float(delta_b) / (float(delta_m) + SCORE_EPS))
        shock.append(0.5 * s_delta + 0.5 * clip01(eta_b / (eta_m + SCORE_EPS)))
    robust = 0.0 if any(bool(c.get("diverged", False)) or bool(c.get("nan_detected", False))
                        or _is(c.get("positivity_ok", True), False) for c in method_cases) else 1.0
    return _scores(_order_score(method_cases), diss, shock, robust)


def tradeoff_ok(method_scores: Mapping[str, float], baseline_scores: Mapping[str, float]) -> Tuple[bool, str]:
    """No order regression and an improvement in dissipation, shock quality or composite
    (Candidate.jl:104-129)."""
    f = lambda d, k: float(d[k])
    order_ok = f(method_scores, "order_preservation") >= f(baseline_scores, "order_preservation") - 1e-12
    improved_diss = f(method_scores, "dissipation") > f(baseline_scores, "dissipation") + 1e-12
    improved_shock = f(method_scores, "shock_quality") > f(baseline_scores, "shock_quality") + 1e-12
    composite_win = f(method_scores, "composite") > f(baseline_scores, "composite") + 1e-12
    notes = ["order_preserved" if order_ok else "order_regressed"]
    if improved_diss:
        notes.append("dissipation_improved")
    if improved_shock:
        notes.append("shock_improved")
    if composite_win:
        notes.append("composite_improved")
    return order_ok and (improved_diss or improved_shock or composite_win), ",".join(notes)


def classify_candidate(method_report: Mapping[str, Any], baseline_report: Mapping[str, Any],
                       delta: float = DEFAULT_SCORE_MARGIN, vtk_produced: bool = False) -> Dict[str, Any]:
    """Candidate status of a method report against a baseline report: rejected, pass_gates,
    promising or accepted_candidate, with the absolute and relative scores (Candidate.jl:136-183)."""
    overall = _is(method_report.get("overall_pass", False), True)
    m_scores = method_report["summary"]["scores"]
    b_scores = baseline_report["summary"]["scores"]
    rel = score_suite_relative(method_report["cases"], baseline_report["cases"])
    c_m, c_b = float(m_scores["composite"]), float(b_scores["composite"])
    margin = c_m - c_b
    t_ok, t_notes = tradeoff_ok(m_scores, b_scores)
    if not overall or method_report.get("diverged", False) or method_report.get("nan_detected", False):
        status = "rejected"
    elif margin >= float(delta) and t_ok:
        status = "accepted_candidate" if vtk_produced else "promising"
    else:
        status = "pass_gates"
    return {
        "candidate_status": status, "baseline_name": baseline_report.get("method_name"),
        "baseline_composite": c_b, "composite": c_m, "composite_margin": margin,
        "score_margin_threshold": float(delta), "vtk_produced": vtk_produced,
        "tradeoff_ok": t_ok, "tradeoff_notes": t_notes, "relative_scores": rel,
        "absolute_scores": m_scores, "baseline_absolute_scores": b_scores,
    }


def collect_hard_gate_failures(cases) -> List[str]:
    """Divergence, failed order gates, 1D Euler positivity and failed conservation gates
    (Scoring.jl:91-109)."""
    fails = []
    for c in cases:
        name = c.get("name", "?")
        if c.get("diverged", False) or c.get("nan_detected", False):
            fails.append(f"{name} diverged/nan")
        if c.get("case_type", "") == "smooth_order" and not c.get("order_pass", True):
            fails.append(f"{name} order_pass failed")
        if c.get("equation", "") == "euler1d" and _is(c.get("positivity_ok", True), False):
            fails.append(f"{name} positivity failed")
        if c.get("conservation_metric", "none") != "none" and not c.get("conservation_pass", True):
            fails.append(f"{name} conservation_pass failed")
    return fails

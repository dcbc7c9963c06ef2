"""Schema v1 verification reports (src/verification/Report.jl, CaseReport.jl, schema_keys.jl):
case dictionaries with the reference's keys, the report skeleton with absolute scores filled
in, the key-contract validator, and JSON write/load."""
from __future__ import annotations

import datetime as _dt
import json
import os
import subprocess
from typing import Any, Dict, Iterable, List, Optional

from .scoring import (DEFAULT_SCORING_WEIGHTS, SCHEMA_VERSION, SCORING_FORMULA_VERSION,
                      score_suite_absolute)
from .state import SchemeConfig

PACKAGE = "FRForge"
PACKAGE_VERSION = "0.1.0"

REQUIRED_TOP_LEVEL_KEYS = (   # schema_keys.jl:15-36
    "schema_version", "package", "package_version", "git_commit", "timestamp_utc",
    "julia_version", "command", "suite", "method_name", "method_params", "baseline_name",
    "overall_pass", "diverged", "nan_detected", "wall_time_sec", "scoring_weights",
    "scoring_formula_version", "hard_gate_failures", "cases", "summary",
)
REQUIRED_SUMMARY_KEYS = ("n_cases", "n_passed", "n_failed", "scores")
REQUIRED_SCORE_KEYS = ("order_preservation", "dissipation", "shock_quality", "robustness", "composite")
REQUIRED_CASE_KEYS = (
    "name", "case_type", "equation", "p", "capturing_method", "pass", "diverged",
    "nan_detected", "conservation_residual", "conservation_pass", "conservation_metric",
    "positivity_ok", "wall_time_sec", "metrics",
)


def case_report_dict(*, name, case_type, equation, p, capturing_method="null", passed,
                     diverged=False, nan_detected=False, conservation_residual=0.0,
                     conservation_pass=True, conservation_metric="none", positivity_ok=True,
                     wall_time_sec=0.0, n_elements=None, t_final=None, excess_dissipation=None,
                     shock_thickness=None, shock_thickness_unit=None, overshoot=None,
                     metrics=None, extra=None) -> Dict[str, Any]:
    """The standard case keys; optional scoring fields are omitted when None
    (CaseReport.jl:13-66).  `passed` is the reference's `pass` keyword."""
    d = {
        "name": str(name), "case_type": str(case_type), "equation": str(equation), "p": int(p),
        "capturing_method": str(capturing_method), "pass": bool(passed),
        "diverged": bool(diverged), "nan_detected": bool(nan_detected),
        "conservation_residual": conservation_residual, "conservation_pass": bool(conservation_pass),
        "conservation_metric": str(conservation_metric), "positivity_ok": bool(positivity_ok),
        "wall_time_sec": float(wall_time_sec),
    }
    opt = {"n_elements": None if n_elements is None else int(n_elements), "t_final": t_final,
           "excess_dissipation": excess_dissipation, "shock_thickness": shock_thickness,
           "shock_thickness_unit": shock_thickness_unit, "overshoot": overshoot}
    d.update({k: v for k, v in opt.items() if v is not None})
    if metrics is not None:
        d["metrics"] = dict(metrics)
    for k, v in (extra or {}).items():
        d[str(k)] = v
    return d


def scheme_dict(scheme: SchemeConfig) -> Dict[str, str]:
    """SchemeConfig.jl:41-50"""
    return {"points": "GL" if scheme.points == "gl" else "GLL",
            "flux": "Rusanov" if scheme.flux == "rusanov" else "HLLC",
            "time": "SSP-RK3" if scheme.time == "ssp_rk3" else "SSP-RK2",
            "points_symbol": scheme.points, "flux_symbol": scheme.flux, "time_symbol": scheme.time}


def git_commit_short() -> str:
    """Short SHA of the working tree, or "unknown" (Report.jl:27-37)."""
    try:
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        out = subprocess.run(["git", "-C", root, "rev-parse", "--short", "HEAD"],
                             capture_output=True, text=True, timeout=10).stdout.strip()
        return out or "unknown"
    except Exception:
        return "unknown"


def empty_scores() -> Dict[str, Any]:
    return {k: None for k in REQUIRED_SCORE_KEYS}


def report_skeleton(*, command="test", suite="smoke", method_name="null", method_params=None,
                    baseline_name=None, overall_pass=True, diverged=False, nan_detected=False,
                    wall_time_sec=0.0, hard_gate_failures=(), cases=(), fill_scores=True,
                    scheme: Optional[SchemeConfig] = None) -> Dict[str, Any]:
    """A schema v1 report with summary counts and absolute scores (Report.jl:60-126).  The
    runtime-version key the schema requires holds this host's description, since the solver
    here is not Julia."""
    cases = list(cases)
    n_passed = sum(1 for c in cases if c.get("pass", False) is True)
    scores = score_suite_absolute(cases) if fill_scores and cases else empty_scores()
    return {
        "schema_version": SCHEMA_VERSION, "package": PACKAGE, "package_version": PACKAGE_VERSION,
        "git_commit": git_commit_short(),
        "timestamp_utc": _dt.datetime.now(_dt.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "julia_version": "none (frforge_amd, HIP/gfx950)",
        "command": str(command), "suite": str(suite), "method_name": str(method_name),
        "method_params": dict(method_params or {}),
        "scheme": scheme_dict(scheme or SchemeConfig()),
        "baseline_name": baseline_name, "overall_pass": bool(overall_pass),
        "diverged": bool(diverged), "nan_detected": bool(nan_detected),
        "wall_time_sec": float(wall_time_sec),
        "scoring_weights": dict(DEFAULT_SCORING_WEIGHTS),
        "scoring_formula_version": SCORING_FORMULA_VERSION,
        "hard_gate_failures": list(hard_gate_failures), "cases": cases,
        "summary": {"n_cases": len(cases), "n_passed": n_passed,
                    "n_failed": len(cases) - n_passed, "scores": scores},
    }


def validate_report_keys(report) -> List[str]:
    """Missing or invalid keys of the schema v1 contract; empty when valid (Report.jl:134-197)."""
    errors: List[str] = []
    if not isinstance(report, dict):
        return ["report is not a dictionary"]
    errors += [f"missing top-level key: {k}" for k in REQUIRED_TOP_LEVEL_KEYS if k not in report]
    if errors:
        return errors
    if report["schema_version"] != SCHEMA_VERSION:
        errors.append(f"schema_version must be {SCHEMA_VERSION}, got {report['schema_version']}")
    if report["scoring_formula_version"] != SCORING_FORMULA_VERSION:
        errors.append(f"scoring_formula_version must be {SCORING_FORMULA_VERSION}, "
                      f"got {report['scoring_formula_version']}")
    if "overall_score" in report:
        errors.append("overall_score is forbidden; use summary.scores.composite only")
    summary = report["summary"]
    if not isinstance(summary, dict):
        return errors + ["summary must be an object"]
    errors += [f"missing summary key: {k}" for k in REQUIRED_SUMMARY_KEYS if k not in summary]
    scores = summary.get("scores")
    if isinstance(scores, dict):
        errors += [f"missing summary.scores key: {k}" for k in REQUIRED_SCORE_KEYS if k not in scores]
    else:
        errors.append("summary.scores must be an object")
    cases = report["cases"]
    if not isinstance(cases, list):
        return errors + ["cases must be an array"]
    for i, case in enumerate(cases, start=1):
        if not isinstance(case, dict):
            errors.append(f"cases[{i}] must be an object")
            continue
        errors += [f"cases[{i}] missing key: {k}" for k in REQUIRED_CASE_KEYS if k not in case]
    return errors


def _json_default(v):
    try:
        import numpy as np
        if isinstance(v, np.generic):
            return v.item()
        if isinstance(v, np.ndarray):
            return v.tolist()
    except ImportError:
        pass
    raise TypeError(f"not JSON serialisable: {type(v)}")


def _finite_json(v):
    """NaN and infinities are written as null: JSON has no literal for them."""
    if isinstance(v, float) and (v != v or v in (float("inf"), float("-inf"))):
        return None
    if isinstance(v, dict):
        return {k: _finite_json(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_finite_json(x) for x in v]
    return v


def write_report(path: str, report: Dict[str, Any]) -> Dict[str, Any]:
    """Validate and write a report as indented JSON (Report.jl:205-216)."""
    errs = validate_report_keys(report)
    if errs:
        raise ValueError("report failed validation: " + "; ".join(errs))
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(_finite_json(json.loads(json.dumps(report, default=_json_default))), f, indent=2,
                  sort_keys=True)
        f.write("\n")
    return report


def load_report(path: str) -> Dict[str, Any]:
    with open(path) as f:
        return json.load(f)

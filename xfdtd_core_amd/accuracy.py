"""Field error of a reduced-precision run against an FP64 run of the same scenario.

Two normalisations are reported for every snapshot frame, per field group
(E = Ex, Ey, Ez together; H = Hx, Hy, Hz together):

* ``own``  -- ||a - b|| / ||b|| with the frame's own norm, the plain relative L2;
* ``peak`` -- ||a - b|| / max over the run's frames of ||b||.

FP32 rounding leaves an error that is nearly constant in absolute terms once
the source has fired (a curl-free, charge-like residue that does not
propagate out through the CPML), while the physical field decays by orders of
magnitude after the pulse has left the box.  The own-norm figure of a
late-time frame therefore measures how far the field has decayed as much as
the arithmetic; the peak-normalised figure measures the arithmetic.  Both are
stated.  Probe time series use the plain relative L2 of the whole series.
"""

from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np

GROUPS = {"E": ("EX", "EY", "EZ"), "H": ("HX", "HY", "HZ")}


def _norm2(a: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    return float(np.dot(a, a))


def snapshot_errors(test: Dict[str, Sequence[np.ndarray]], ref: Dict[str, Sequence[np.ndarray]],
                    steps: Sequence[int] = None) -> List[dict]:
    """`test` / `ref` map a field name to its list of snapshot frames.

    Returns one entry per frame: {"frame": f, "step": s, "E": {"own", "peak"},
    "H": {...}}.  A figure whose reference norm is zero (a frame before the
    source has fired) is None: there is nothing to be relative to."""
    nframes = min(len(ref[f]) for g in GROUPS.values() for f in g if f in ref)
    out = []
    peaks = {}
    for g, fields in GROUPS.items():
        fs = [f for f in fields if f in ref]
        peaks[g] = max(np.sqrt(sum(_norm2(ref[f][i]) for f in fs)) for i in range(nframes)) if fs else 0.0
    for i in range(nframes):
        row = {"frame": i}
        if steps is not None:
            row["step"] = int(steps[i])
        for g, fields in GROUPS.items():
            fs = [f for f in fields if f in ref]
            if not fs:
                continue
            d = np.sqrt(sum(_norm2(np.asarray(test[f][i], np.float64) - np.asarray(ref[f][i], np.float64))
                            for f in fs))
            n = np.sqrt(sum(_norm2(ref[f][i]) for f in fs))
            own = float(d / n) if n > 0 else None
            peak = float(d / peaks[g]) if peaks[g] > 0 else None
            row[g] = {"own": own, "peak": peak}
        out.append(row)
    return out


def series_error(test: np.ndarray, ref: np.ndarray):
    """Relative L2 of a probe's whole time series; None when the reference
    series is all zero (the probe never saw the field)."""
    t = np.asarray(test, np.float64).ravel()
    r = np.asarray(ref, np.float64).ravel()
    n = np.sqrt(_norm2(r))
    d = np.sqrt(_norm2(t - r))
    return float(d / n) if n > 0 else None

"""Comparison of dumps against reference vectors.

Restates tools/validate.py:20-67 (maximum and mean absolute difference,
location of the maximum, non-zero exit on a length mismatch) and adds what the
reference leaves out: a tolerance that fails the comparison (the reference
never fails on values, tools/validate.py:38-46).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Report:
    current: str
    reference: str
    maximum: float
    average: float
    location: tuple
    passed: bool

    def text(self):
        lines = ["**************************************************", "Validation:",
                 f"  Comparison Matrix : {self.current}", f"  Reference Matrix  : {self.reference}",
                 f"  Maximum Delta     : {self.maximum}"]
        if self.maximum > 1e-5:
            lines.append(f"  Maximum Location  : {self.location}")
        lines.append(f"  Average Delta     : {self.average}")
        lines.append(f"  Result            : {'PASS' if self.passed else 'FAIL'}")
        return "\n".join(lines)


class LengthMismatch(ValueError):
    pass


def load_values(path):
    return np.loadtxt(path, ndmin=1).ravel()


def compare(current_path, reference_path, atol=1e-5, rtol=1e-5):
    """Max/mean |current - reference|; passes when every element satisfies
    |d| <= atol + rtol * |reference|.  Location as the reference reports it:
    (row // 3, row % 3) = (fiber or coefficient triple, component)."""
    cur, ref = load_values(current_path), load_values(reference_path)
    if cur.shape != ref.shape:
        raise LengthMismatch(f"{current_path}: {cur.size} values, reference {ref.size}")
    d = np.abs(cur - ref)
    k = int(np.argmax(d)) if d.size else 0
    passed = bool(np.all(d <= atol + rtol * np.abs(ref)))
    return Report(current_path, reference_path, float(d.max(initial=0.0)), float(d.mean()) if d.size else 0.0,
                  (k // 3, k % 3), passed)

"""Shared helpers of the tests: the seeded workloads and error measures."""
import math

import numpy as np

from processseg_amd.workloads import (big_flow as config2_flow, big_panel as config2_panel,  # noqa: F401
                                      plain_flow as config1_flow, plain_panel as config1_panel)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    r = math.sqrt(float(np.sum(b * b)))
    d = math.sqrt(float(np.sum((a - b) ** 2)))
    return d / r if r > 0 else d


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))

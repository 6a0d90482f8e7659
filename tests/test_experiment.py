"""The experiment runner's log line matches the reference's format string (main.cu:56-60), so logs written by
either can be parsed the same way. CPU only."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

# The format string of glog in the reference's main.cu (runExperiment), applied with C printf semantics.
REFERENCE_FORMAT = ("{-%.3e/+%.3e batchf, %03d batchi, %03d threads} -> {%09.1fms, %09.1fms init, %09.1fms mark, "
                    "%09.1fms comp, %03d iter, %.2e err} %s")


def test_glog_matches_reference_format():
    from pagerank_amd.experiment import glog
    ref = np.array([0.25, 0.25, 0.5])
    ans = SimpleNamespace(ranks=np.array([0.25, 0.26, 0.49]), time=12.34567, initializationTime=0.5,
                          markingTime=0.0625, computationTime=11.75, iterations=7)
    got = glog(ans, ref, "pagerankPruneDynamicFrontierHip", 1, 0.0, 1e-4, 3)
    want = REFERENCE_FORMAT % (0.0, 1e-4, 3, 1, 12.34567, 0.5, 0.0625, 11.75, 7, 0.02,
                               "pagerankPruneDynamicFrontierHip")
    assert got == want

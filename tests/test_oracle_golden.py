"""The FP64 oracle reproduces the reference's printed golden values
(test/Lik_*.cx.exp)."""

import math

import pytest

from oracle.oracle import OracleLik
from tests.golden_util import run_case, tol


def test_oracle_golden(golden):
    fails = []
    for case in golden:
        approx = case.get("approx_tree", False)
        for ev, lnl, sites in run_case(case, OracleLik):
            if "lnL_below" in ev:
                if not lnl < ev["lnL_below"]:
                    fails.append((case["name"], lnl))
                continue
            if "sig" in ev:
                t = 0.5 * 10.0 ** (math.floor(math.log10(abs(ev["lnL"]))) - ev["sig"] + 1) * 1.0001
            else:
                t = tol(ev, approx=approx)
            if abs(lnl - ev["lnL"]) > t:
                fails.append((case["name"], lnl, ev["lnL"], t))
            if sites is not None:
                ts = tol(ev, "site_dec", approx)
                for a, b in zip(sites, ev["sites"]):
                    if abs(a - b) > ts:
                        fails.append((case["name"], "site", a, b))
                assert len(sites) == len(ev["sites"])
    assert not fails, fails

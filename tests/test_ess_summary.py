"""The learned-vs-fixed comparison's figures (sck_amd.ess.summarize_proposals) on synthetic per-proposal runs
whose answers are computed by hand: two temperature groups, fixed widths 1 and 2, the uniform redraw, the
learned width and its finer table. CPU only; the GPU runs that produce such results are in test_learned.py and
the bench's ESS object."""

import math

import numpy as np
import pytest

from sck_amd.ess import summarize_proposals


def _run(seconds, tau_E, tau_M=None, E=(1.0, 2.0)):
    tau_M = tau_M or tau_E
    obs = [dict(T=float(j), E=E[j], E_err=0.1, M=0.5, M_err=0.05, Cv=1.0, Cv_err=0.2, chi=2.0, chi_err=0.3)
           for j in range(2)]
    # ESS/s summed over the groups: 100 chains x 400 samples / (2 tau) per group, over the run's seconds
    ess = lambda tau: sum(100 * 400 / (2.0 * tau[j]) for j in range(2)) / seconds
    return dict(seconds=seconds, tau_E=dict(enumerate(tau_E)), tau_M=dict(enumerate(tau_M)),
                ess_E_per_s=ess(tau_E), ess_M_per_s=ess(tau_M), observables=obs)


def _result():
    return {"fixed_width": 1.5,                                   # the width the training objective picked
            "uniform": _run(1.0, (4.0, 2.0)),
            "fixed_1": _run(1.0, (2.0, 4.0)),
            "fixed_2": _run(1.0, (3.0, 3.0)),
            "learned": _run(1.25, (1.5, 2.0), E=(1.3, 2.0)),
            "learned_fine": _run(1.3, (1.5, 1.6))}


def test_summary_figures_match_hand_computed_values():
    r = summarize_proposals(_result(), 2, (1.0, 2.0))
    # geometric-mean ESS/s of E: width 1 -> sqrt(0.5 x 0.25) = 0.354, width 2 -> 1/3: width 1 is the baseline
    assert r["fixed_width"] == 1.0 and r["fixed_width_by_objective"] == 1.5
    assert r["fixed"] is r["fixed_1"]
    assert math.isclose(r["fixed_scan_gm_ess"]["fixed_1"], math.sqrt(0.125))
    # learned over width 1, per group (tau_a / tau_b)(t_a / t_b): 2/1.5 x 0.8 and 4/2 x 0.8
    assert np.allclose(r["gain_E_per_T"], [16 / 15, 1.6])
    assert math.isclose(r["gain_E"], math.sqrt(16 / 15 * 1.6))
    # the per-group best constant proposal: width 1 at group 0 (0.5), the uniform redraw at group 1 (0.5)
    assert r["per_T_best_fixed"] == ["fixed_1", "uniform"]
    assert math.isclose(r["gain_E_per_T_best_over_fixed"], math.sqrt(2.0))
    assert math.isclose(r["gain_E_vs_per_T_best"], math.sqrt(16 / 15 * 0.8))
    # the table against the finer one: tau ratios 1.5/1.5 and 2/1.6
    assert math.isclose(r["table_vs_fine_tau_E"], math.sqrt(1.25))
    # over the uniform redraw: 4/1.5 x 0.8 and 2/2 x 0.8
    assert math.isclose(r["vs_uniform"]["gain_E"], math.sqrt(32 / 15 * 0.8))
    # the one planted difference: E at group 0, 0.3 apart with errors 0.1 each, against both baselines
    assert math.isclose(r["max_z"], 0.3 / math.hypot(0.1, 0.1))
    assert math.isclose(r["frac_within_2sigma"], 14 / 16)


def test_summary_without_the_fine_table():
    res = _result()
    del res["learned_fine"]
    r = summarize_proposals(res, 2, (1.0, 2.0), fine_table=False)
    assert "table_vs_fine_tau_E" not in r and r["fixed_width"] == 1.0


def test_summary_needs_every_scanned_width():
    res = _result()
    del res["fixed_2"]
    with pytest.raises(KeyError):
        summarize_proposals(res, 2, (1.0, 2.0))

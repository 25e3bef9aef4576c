"""Solute ledger on the GPU (rs_solute_ledger) against the oracle's: a
256^2 coupled-T run with a pull-back every 150 steps (600 steps). Every
path must close its own ledger (the unaccounted change within 1e-13 of the
mass: the interior faces cancel, what remains is the rounding of the c
updates); the FP64 path's terms equal the oracle's (pull-backs 1e-12, the
boundary-face gains 1e-9 relative); the FP32 and surrogate paths' terms are
reported against the oracle's; a batch books its runs bitwise like the
runs stepped alone."""
import sys

import numpy as np
import pytest

from rsolid import RsError
from rsolid.abi import RS_ERR_STATE, RS_PREC_MIXED, RS_PREC_SURROGATE

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-ledger]", *a, flush=True)
    sys.stdout.flush()


def _run(ctx, steps=600, spb=150):
    k = ctx.constants()
    k.step_pull_back = spb
    ctx.set_constants(k)
    ctx.solute_ledger(reset=True)  # starts the GPU ledger at the initial fields
    ctx.step(steps)
    return ctx.solute_ledger()


@pytest.fixture(scope="module")
def oracle_ledger(rs, oracle_mod):
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(7)
    L = _run(st)
    st.close()
    say("oracle", L)
    return L


@pytest.mark.parametrize("prec", [0, RS_PREC_MIXED, RS_PREC_SURROGATE])
def test_ledger_closes_and_matches_the_oracle(rs, oracle_ledger, prec):
    p = rs.config_params(256, True, precision=prec)
    ctx = rs.Context(p)
    ctx.init_config(7)
    if prec == RS_PREC_SURROGATE:
        ctx.surrogate_train()
    L = _run(ctx)
    ctx.close()
    O = oracle_ledger
    say("path", prec, L)
    assert abs(L["unaccounted"]) <= 1e-13 * L["mass"]
    for key in ("pull_back", "edge_x", "edge_y"):
        say(" ", key, "rel diff vs oracle", (L[key] - O[key]) / abs(O[key]))
    if prec == 0:
        assert L["mass0"] == pytest.approx(O["mass0"], rel=1e-14)
        # a difference of two interior-mass sums whose rounding differs
        # (tree vs sequential order): bounded on the scale of the mass
        assert abs(L["pull_back"] - O["pull_back"]) <= 1e-13 * O["mass"]
        assert L["edge_x"] == pytest.approx(O["edge_x"], rel=1e-9)
        assert L["edge_y"] == pytest.approx(O["edge_y"], rel=1e-9)
    else:
        # the FP32 fields (and the network's fit) move the columns a
        # pull-back drops and the near-edge fluxes: the terms agree to the
        # path's field accuracy, while each path's own ledger closes
        tol = 1e-6 if prec == RS_PREC_MIXED else 1e-3
        assert L["pull_back"] == pytest.approx(O["pull_back"], rel=tol)
        assert L["edge_y"] == pytest.approx(O["edge_y"], rel=1e-2)
        say("  raw drift vs oracle", (L["mass"] - O["mass"]) / O["mass"],
            "unaccounted vs oracle", (L["unaccounted"] - O["unaccounted"]) / O["mass"])


def test_batch_books_runs_like_single_runs(rs):
    specs = [(256, 43.24, 0.0842), (320, 30.0, 0.05), (256, 25.0, 0.2)]
    params = [rs.ensemble_params(n, d, v, RS_PREC_MIXED) for n, d, v in specs]
    seeds = [61, 62, 63]
    b = rs.Batch(params)
    for run, s in zip(b.runs, seeds):
        run.init_config(s)
        k = run.constants()
        k.step_pull_back = 100
        run.set_constants(k)
        run.solute_ledger(reset=True)
    b.step(300)
    got = [r.solute_ledger() for r in b.runs]
    b.close()
    for p, s, g in zip(params, seeds, got):
        c = rs.Context(p)
        c.init_config(s)
        k = c.constants()
        k.step_pull_back = 100
        c.set_constants(k)
        c.solute_ledger(reset=True)
        c.step(300)
        want = c.solute_ledger()
        c.close()
        say("batch run", g)
        assert g == want
        assert abs(g["unaccounted"]) <= 1e-13 * g["mass"]


def test_ledger_is_off_until_started(rs):
    """A read before the first reset is refused; a batch whose runs start
    their ledgers between segments records from then on (the batch takes
    the recording kernel), and a run left off keeps refusing."""
    params = [rs.ensemble_params(256, 43.24, 0.0842, RS_PREC_MIXED) for _ in range(2)]
    b = rs.Batch(params)
    for run, s in zip(b.runs, (5, 6)):
        run.init_config(s)
    with pytest.raises(RsError) as e:
        b.runs[0].solute_ledger()
    assert e.value.code == RS_ERR_STATE
    b.step(50)
    b.runs[0].solute_ledger(reset=True)
    b.step(200)
    L = b.runs[0].solute_ledger()
    say("started mid-run", L)
    assert abs(L["unaccounted"]) <= 1e-13 * L["mass"]
    assert L["edge_x"] != 0.0
    with pytest.raises(RsError):
        b.runs[1].solute_ledger()
    b.close()

"""JSON configs of pmp_params for the tools (bench.py --params,
profile_run --params): a base config plus field overrides, enum names,
round trip, and errors for unknown keys and wrong types."""
import json

import pytest

from patchmatch_amd.params import (AGG_MAX, EDGE_ZERO, config0_params, config1_params, params_from_json,
                                   params_to_json)


def test_base_and_overrides(tmp_path):
    p = params_from_json({"base": "config1", "iterations": 4, "agg_mode": "max", "edge_mode": EDGE_ZERO})
    want = config1_params(iterations=4, agg_mode=AGG_MAX, edge_mode=EDGE_ZERO)
    assert p.as_dict() == want.as_dict()
    # default base is config 1; a file and a JSON text give the same
    f = tmp_path / "p.json"
    f.write_text(json.dumps({"seed": 7}))
    assert params_from_json(str(f)).as_dict() == config1_params(seed=7).as_dict()
    assert params_from_json('{"base": "config0"}').as_dict() == config0_params().as_dict()


def test_round_trip():
    p = config0_params(min_patch_sd=2.5, median_rel_thr=0.07, seed=2 ** 40 + 3, sigma_color=33.3)
    assert params_from_json(params_to_json(p)).as_dict() == p.as_dict()


@pytest.mark.parametrize("bad", [{"foo": 1}, {"iterations": 1.5}, {"iterations": "4"}, {"agg_mode": "bogus"},
                                 {"base": "config9"}, {"seed": True}, "[1, 2]", "/nonexistent/params.json"])
def test_errors(bad):
    with pytest.raises(ValueError):
        params_from_json(bad)

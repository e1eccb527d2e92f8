"""Flame parsing and optimisation: the product's C++ host against the oracle's
independent Python model, and both against closed forms.

Pins: cumulative weights are the running sums of w_i / sum(w) with the last
set to 1.0 (flame.hpp:44-60); xforms of weight 0 are dropped and the rest
sorted by libstdc++ std::sort with a.weight > b.weight (flame.hpp:63-83);
error texts are the reference's (flame.hpp, xform.hpp, affine.hpp,
point.hpp, variations.hpp).
"""

import glob
import json
import os

import numpy as np
import pytest

import flames as TF
from conftest import FLAMES, read_flame


def both(ffr_mod, oracle_mod, text, precision="fp64"):
    got = ffr_mod.check_flame(text, precision)
    mode = oracle_mod.F64_U64 if precision == "fp64" else oracle_mod.F32_U64
    f = oracle_mod.OracleFlame(text, mode)
    return got, (f.xform_order, f.cum_weights)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_product_and_oracle_agree_on_test_flames(ffr_mod, oracle_mod, precision):
    texts = [TF.fern(), TF.rotations(), TF.tetrahedron(), TF.one_d(), TF.tied_weights(),
             TF.color_flame(3), TF.occasional_bad(), read_flame("sierpinski_1024.json"),
             read_flame("random_color_3840x2160.json"), read_flame("tetrahedron_512.json")]
    texts += [open(p).read() for p in sorted(glob.glob(os.path.join(FLAMES, "batch", "*.json")))[:64]]
    for t in texts:
        got, want = both(ffr_mod, oracle_mod, t, precision)
        assert got[0] == want[0]
        assert got[1] == want[1]


def test_cumulative_weights_closed_form(ffr_mod):
    t = json.loads(TF.fern())
    for x, w in zip(t["xforms"], [1.0, 1.0, 1.0, 3.0]):
        x["weight"] = w
    ids, cw = ffr_mod.check_flame(json.dumps(t))
    assert ids == [3, 0, 1, 2]
    assert cw == [3.0 / 6.0, 3.0 / 6.0 + 1.0 / 6.0, 3.0 / 6.0 + 1.0 / 6.0 + 1.0 / 6.0, 1.0]


def test_zero_weights_dropped(ffr_mod, oracle_mod):
    t = json.loads(TF.fern())
    t["xforms"][1]["weight"] = 0
    ids, _ = ffr_mod.check_flame(json.dumps(t))
    assert 1 not in ids and len(ids) == 3
    t["xforms"] = [dict(t["xforms"][0], weight=0)]
    with pytest.raises(ffr_mod.FfrError, match="no xforms remaining after optimization"):
        ffr_mod.check_flame(json.dumps(t))


@pytest.mark.parametrize("n", [3, 16, 17, 40, 200])
def test_sort_order_ties_match_libstdcxx(ffr_mod, oracle_mod, n):
    """More than 16 tied weights leave std::sort's insertion-sort path; the
    product calls libstdc++ std::sort, the oracle restates its introsort."""
    rng = np.random.default_rng(n)
    for trial in range(20):
        w = [float(v) for v in rng.choice([0.25, 0.5, 1.0, 2.0], size=n)]
        text = json.dumps({"dimensions": 1, "size": [8], "bounds": [[-1, 1]],
                           "xforms": [{"weight": x, "variations": [{"name": "linear", "weight": 1}]}
                                      for x in w]})
        ids, _ = ffr_mod.check_flame(text)
        assert ids == oracle_mod.sort_order(w)
        # decreasing weight order, whatever the tie order
        assert all(w[a] >= w[b] for a, b in zip(ids, ids[1:]))


def test_sort_order_small_is_stable(oracle_mod):
    # <= 16 elements: insertion sort, ties keep file order
    assert oracle_mod.sort_order([1, 2, 1, 2, 1]) == [1, 3, 0, 2, 4]


def _bad(mutate):
    doc = json.loads(TF.color_flame(3))
    mutate(doc)
    return json.dumps(doc)


BAD_FLAMES = [
    (lambda d: d.pop("size"), "Flame(): Json::operator[](char*): key does not exist: size"),
    (lambda d: d.update(size=[10]), "Flame(): incorrect size length"),
    (lambda d: d.update(size=[0, 10]), "Flame(): size[0] out of range"),
    (lambda d: d.update(size=[10, 70000]), "Flame(): size[1] out of range"),
    (lambda d: d.update(size=["a", 10]), "Flame(): error parsing size[0]: Json::floatValue(): type is not float"),
    (lambda d: d.update(bounds=[[1, 0], [0, 1]]), "Flame(): bounds[0] low >= high"),
    (lambda d: d.update(bounds=[[0, 1], [0, 1e11]]), "Flame(): bounds[1] out of range"),
    (lambda d: d.update(bounds=[[0, 1, 2], [0, 1]]), "Flame(): bounds[0] wrong format"),
    (lambda d: d.update(color_dimensions=200), "Flame(): too many color dimensions"),
    (lambda d: d.update(color_speed=1.5), "Flame(): color speed out of range"),
    (lambda d: d.update(xforms=[]), "Flame(): no xforms"),
    (lambda d: d.update(xforms=5), "Flame(): cannot parse xforms: Json::arrayValue(): type is not array"),
    (lambda d: d["xforms"][0].update(weight=-1), "Flame(): error parsing xforms[0]: XForm(): weight is negative"),
    (lambda d: d["xforms"][1].pop("weight"),
     "Flame(): error parsing xforms[1]: XForm(): cannot parse weight: Json::operator[](char*): key does not exist: weight"),
    (lambda d: d["xforms"][0].update(pre_affine=[1]),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse pre_affineAffine(Json&): not an object"),
    (lambda d: d["xforms"][0].update(pre_affine={"A": [[1, 0]]}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse pre_affineAffine(Json&): A is wrong size"),
    (lambda d: d["xforms"][0].update(post_affine={"b": [1, 2, 3]}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse post_affineAffine(Json&): error parsing b: Point(Json&): incorrect array size: 3"),
    (lambda d: d["xforms"][0].update(pre_affine={"A": [[1, "x"], [0, 1]]}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse pre_affineAffine(Json&): error parsing A[0]: Point(Json&): entry is not a number"),
    (lambda d: d["xforms"][0]["variations"].append({"name": "nope", "weight": 1}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse variations: unknown variation: nope"),
    (lambda d: d["xforms"][0]["variations"].append({"name": "disc2", "weight": 1, "rotation": 1}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse variations: Json::operator[](char*): key does not exist: twist"),
    (lambda d: d["xforms"][0]["variations"].append({"name": "boarders", "weight": 1, "prob": 2}),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse variations: boarders probability out of range"),
    (lambda d: d["xforms"][0].update(color=[0.5, 0.5]),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse color: color length incorrect"),
    (lambda d: d["xforms"][0].update(color=[0.5, 0.5, 1.5]),
     "Flame(): error parsing xforms[0]: XForm(): cannot parse color: color coordinate out of range"),
    (lambda d: d["xforms"][0].update(color=0.5),
     "Flame(): error parsing xforms[0]: Json::arrayValue(): type is not array"),
    (lambda d: d["xforms"][0].update(color_speed=-0.1),
     "Flame(): error parsing xforms[0]: XForm(): color speed issue: color speed out of range"),
    (lambda d: d["final_xform"].pop("variations"),
     "Flame(): error parsing final xform: XForm(): cannot parse variations: Json::operator[](char*): key does not exist: variations"),
]


@pytest.mark.parametrize("mutate,message", BAD_FLAMES)
def test_error_messages_match_reference_texts(ffr_mod, oracle_mod, mutate, message):
    text = _bad(mutate)
    with pytest.raises(ffr_mod.FfrError) as e:
        ffr_mod.check_flame(text)
    assert str(e.value) == message
    import oracle.model as M
    with pytest.raises(M.FlameError) as e2:
        M.parse_flame(text)
    assert str(e2.value) == message


def test_three_d_axes_errors(ffr_mod):
    doc = json.loads(TF.tetrahedron())
    doc["xforms"][0]["variations"] = [{"name": "swirl", "weight": 1, "axis_x": 0}]
    with pytest.raises(ffr_mod.FfrError, match="key does not exist: axis_y"):
        ffr_mod.check_flame(json.dumps(doc))
    doc["xforms"][0]["variations"] = [{"name": "swirl", "weight": 1, "axis_x": 1, "axis_y": 1}]
    with pytest.raises(ffr_mod.FfrError, match="axes are not distinct"):
        ffr_mod.check_flame(json.dumps(doc))
    doc["xforms"][0]["variations"] = [{"name": "swirl", "weight": 1, "axis_x": 0, "axis_y": 3}]
    with pytest.raises(ffr_mod.FfrError, match="axis_y index out of range"):
        ffr_mod.check_flame(json.dumps(doc))


def test_one_d_rejects_2d_variations(ffr_mod):
    doc = json.loads(TF.one_d())
    doc["xforms"][0]["variations"] = [{"name": "swirl", "weight": 1}]
    with pytest.raises(ffr_mod.FfrError, match="unknown variation: swirl"):
        ffr_mod.check_flame(json.dumps(doc))


def test_json_comments_and_numbers(ffr_mod):
    text = """// leading comment
    { "dimensions": 2, /* block */ "size": [4, 4], "bounds": [[-1e0, 1.0], [-1, 1]],
      "xforms": [ {"weight": 1, "variations": [{"name": "linear", "weight": 1}]} ] }
    """
    ids, cw = ffr_mod.check_flame(text)
    assert ids == [0] and cw == [1.0]
    with pytest.raises(ffr_mod.FfrError, match="parse error"):
        ffr_mod.check_flame(text + "}")

"""The .geom scene-language reader of the host layer against the reference's own parser.

The expected scenes are the reference parser's output (tests/golden/make_golden.py runs the
reference's meep-free sources on each script): context numbers, sources, monitor clusters,
parse error code, and the resolved CSG tree and metadata of every root, compared exactly.

* geom/quirks.geom and geom/small_run.geom are scripts of this repository, written so that a
  different reading of an evaluation rule (operator splitting, ternary binding, '=' inside
  comparisons, undefined names, whitespace fusion in lists, trailing commas in geometry lists,
  literal instances, the Difference/Complement/Rotate trees) changes a number.
* The reference's own junction and test scripts are read from the reference tree when it is
  present (they are not copied here).
"""
import json
import os

import pytest

from sim_juncs_amd import host

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = os.environ.get("SJF_REFERENCE", "/root/reference")


def load_golden(name):
    with open(os.path.join(G, f"scene_{name}.json")) as f:
        return json.load(f)


def run_loader(golden, geom_dir):
    st = golden["settings"]
    cwd = os.getcwd()
    os.chdir(geom_dir)
    try:
        return host.geom_scene(os.path.join(geom_dir, st["geom_fname"]), st["pml_thickness"], st["len"],
                               st["um_scale"], st["out_dir"], st["user_opts"])
    finally:
        os.chdir(cwd)


def assert_same_scene(mine, er, golden):
    assert er == golden["parse_error"]
    for key in ("context_numbers", "sources", "monitor_clusters", "roots"):
        assert mine[key] == golden[key], key


@pytest.mark.parametrize("name", ["quirks", "small_run"])
def test_own_scripts_match_reference_parser(name):
    g = load_golden(name)
    mine, er = run_loader(g, os.path.join(G, "geom"))
    assert_same_scene(mine, er, g)


def test_evaluation_rules_known_answers():
    """The rules, as numbers that follow from them by hand (and that the reference gives)."""
    g = load_golden("quirks")
    mine, _ = run_loader(g, os.path.join(G, "geom"))
    v = dict(mine["context_numbers"])
    assert v["chain_sub"] == 11          # 10 - (2 - 3): the first '-' splits
    assert v["chain_div"] == 4           # 8 / (4 / 2)
    assert v["power"] == 512             # 2 ^ (3 ^ 2)
    assert v["tern_plus"] == 8           # 1 + ((0) ? 5 : 7)
    assert v["tern_else"] == 5           # ((0) ? 1 : 2) + ((1) ? 3 : 4)
    assert v["cmp_eq"] == 2              # every '=' of a statement starts its right-hand side anew
    assert v["cmp_le"] == 1
    assert v["unknown_plus"] == 1        # an undefined name is 0
    assert v["neg_name"] == 0            # "-name" is not a number and no name
    assert v["neg_num"] == -1.5
    assert v["expo"] == 250.001          # no split after an 'e'
    assert v["twice"] == 8 and v["extra"] == 4   # --opts statements run first
    assert v["scaled"] == 3              # um_to_l with l_per_um = 2
    # whitespace between numbers fuses them: "1.0 2.0 3.0" -> "1.02.03.0" -> 1.02
    meta = [r["metadata"] for r in mine["roots"]]
    assert meta[1]["susceptibilities"][0] == [1.02, "lorentz"]
    # roots come newest first; a named combine_type is metadata too
    assert meta[0] == {"eps": 1.5, "combine_type": "intersect"}
    # a CW source without an end time runs forever
    assert mine["sources"][2]["end_time"] == 1.7976931348623157e308


def test_trailing_comma_shapes_the_tree():
    """A trailing comma in a two-entry geometry list adds a third entry (the number 0) that is
    skipped, so the second box is no longer last: it is chained into a nested node."""
    g = load_golden("small_run")
    mine, _ = run_loader(g, os.path.join(G, "geom"))
    t = mine["roots"][0]["tree"]
    assert t["children"][0]["kind"] == "box"
    nested = t["children"][1]
    assert nested["kind"] == "composite" and nested["children"][0]["kind"] == "box"
    assert nested["children"][1] is None


@pytest.mark.parametrize("name", ["test", "run", "span", "box", "bowtie", "sidetie", "bowtie_inf"])
def test_reference_scripts_match_reference_parser(name):
    g = load_golden(name)
    d = os.path.join(REF, g["case"]["dir"])
    if not os.path.exists(os.path.join(d, g["settings"]["geom_fname"])):
        pytest.skip("the reference tree is not present")
    mine, er = run_loader(g, d)
    assert_same_scene(mine, er, g)


def test_missing_script_raises():
    with pytest.raises(host.HostError):
        host.geom_scene("/nonexistent/scene.geom", 1.0, 2.0)


def test_error_stops_the_script_and_drops_the_roots(tmp_path):
    """An unknown function is error 6 (E_BAD_TYPE, as the reference's sidetie script shows):
    what ran before it stays (numbers, sources), nothing after it runs and no roots are built."""
    geom = tmp_path / "broken.geom"
    geom.write_text(
        "a = 2\n"
        "Gaussian_source(\"Ey\", 0.8, 1.0, 2.0, 0.0, Box([0,0,1], [length,length,1]))\n"
        "Composite(eps = 3.0, [Box([0,0,0], [1,1,1])])\n"
        "b = not_a_function(1)\n"
        "c = 3\n")
    scene, er = host.geom_scene(str(geom), 1.0, 2.0)
    assert er == 6
    names = [n for n, _ in scene["context_numbers"]]
    assert "a" in names and "b" not in names and "c" not in names
    assert len(scene["sources"]) == 1 and scene["sources"][0]["component"] == 1
    assert scene["roots"] == []

"""The YAML-driven caller on the CPU: the binning-text reader (against the
reference's own binning files and the parser rules of BinSet.cpp:143-257),
the selection-formula evaluator, the event store, and the assembly of the
cache tables from a configuration, checked against an independent
event-by-event construction.  No device work: binning goes through the
oracle here."""
import os

import numpy as np
import pytest
import yaml

from gundam_amd import _native as N
from gundam_amd.binning import BinningError, parse_binning_text, read_binning_text
from gundam_amd.eventstore import EventStore, StoreError, write_store
from gundam_amd.formula import FormulaError, evaluate, select
from gundam_amd.propagator import ConfigError, Propagator
from oracle import pyoracle

import caller_fixture

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "binning")


def test_reference_binning_files():
    bs = read_binning_text(os.path.join(GOLD, "200CovarianceFit-binning.txt"))
    assert bs.n_bins == 36 and bs.variables() == ["A", "B"]
    assert bs.bins[0].edges == {"A": (-10.0, -2.0), "B": (-10.0, -2.0)}
    assert bs.bins[35].edges == {"A": (2.0, 10.0), "B": (2.0, 10.0)}
    bc = read_binning_text(os.path.join(GOLD, "200MultiParFitBicubic-binning.txt"))
    assert bc.n_bins == 16  # blank lines between blocks are skipped
    assert bc.bins[4].edges == {"A": (-1.0, 0.0), "B": (-10.0, -1.0)}


def test_binning_parser_rules():
    text = """# comment line
      # indented comment
    variables: sample p p cos cos   # trailing comment
    0  0 1  -1 0
    1  0.5 2 0 1
    variables: q q
    3 4
    """
    bs = parse_binning_text(text)
    assert bs.n_bins == 3
    assert bs.bins[0].edges == {"cos": (-1.0, 0.0), "p": (0.0, 1.0), "sample": (0.0, 0.0)}
    assert list(bs.bins[0].edges) == ["cos", "p", "sample"]  # sorted by name
    assert bs.bins[2].edges == {"q": (3.0, 4.0)}
    assert bs.variables() == ["cos", "p", "sample", "q"]
    e = bs.edge_array(["p", "q"])
    assert e[0, 1, 0] == -np.inf and e[0, 1, 1] == np.inf and tuple(e[2, 1]) == (3.0, 4.0)
    for bad in ("1 2\n", "variables: a a a\n1 2 3\n", "variables: a b a\n1 2 3\n",
                "variables: a a\n1\n"):
        with pytest.raises(BinningError):
            parse_binning_text(bad)


def test_binning_semantics_against_bin_definition():
    """First bin whose every edge holds the value: [min, max), or equality
    when min == max (Bin.cpp:167-176), via the oracle's binning."""
    bs = parse_binning_text("variables: s a a\n0 0 1\n0 1 2\n1 0 2\n")
    vals = np.array([[0, 0.0], [0, 1.0], [0, 2.0], [1, 1.5], [2, 0.5], [0, -0.1]], float)
    got = pyoracle.bin_events(vals[:, [1, 0]], bs.edge_array(["a", "s"]))
    assert list(got) == [0, 1, -1, 2, -1, -1]


def test_formula_evaluator():
    cols = {"A": np.array([1.0, -2.0, 3.0]), "B": np.array([0.0, 1.0, 2.0]),
            "x.y": np.array([5.0, 6.0, 7.0])}
    col = cols.__getitem__
    np.testing.assert_array_equal(evaluate("[A] * 2 + B", col, 3), [2.0, -3.0, 8.0])
    np.testing.assert_array_equal(select("[A] > 0 && !([B] > 1.5)", col, 3), [True, False, False])
    np.testing.assert_array_equal(select("[A] < 0 || [B] == 2", col, 3), [False, True, True])
    np.testing.assert_array_equal(select("0 < [B] <= 1", col, 3), [False, True, False])
    np.testing.assert_array_equal(evaluate("abs([A]) + pow([B], 2) + [x.y]", col, 3),
                                  [6.0, 9.0, 14.0])
    assert select("", col, 3).all() and select(None, col, 3).all()
    for bad in ("__import__('os')", "[A].real", "'s'", "open(1)", "[A] if 1 else 2"):
        with pytest.raises(FormulaError):
            evaluate(bad, col, 3)


def test_event_store_round_trip(tmp_path):
    p = str(tmp_path / "s.store")
    st = write_store(p, {"a": [1.0, 2.0, 3.0]}, [0.5, 1.0, 2.0],
                     {"sp": ("monotonic", [np.arange(9.0), np.zeros(0), np.ones(9)])})
    st = EventStore(p)
    assert st.n_events == 3 and list(st.column("a")) == [1.0, 2.0, 3.0]
    t, beg, data = st.dial_leaf("sp")
    assert t == N.DIAL_MONOTONIC and list(beg) == [0, 9, 9, 18] and data[17] == 1.0
    with pytest.raises(StoreError):
        st.column("nope")
    with pytest.raises(StoreError):
        write_store(str(tmp_path / "t"), {"a": [1.0]}, [1.0, 2.0])
    with pytest.raises(StoreError):
        EventStore(str(tmp_path))


def _expected_tables(prop):
    """Event-by-event restatement of the configuration semantics."""
    cfg_bins = [s.binning for s in prop.samples]
    base, bins, types, inputs, payload, counts, inputs2 = [], [], [], [], [], [], []
    for s, bs in zip(prop.samples, cfg_bins):
        for ds in prop.datasets:
            if s.datasets is not None and ds["name"] not in s.datasets:
                continue
            for st in ds["mc"]:
                col = st.column
                n = st.n_events
                cut = select(ds["mc_cut"], col, n) & select(s.selection, col, n)
                wf = evaluate(ds["mc_weight"], col, n) if ds["mc_weight"] else np.ones(n)
                for e in range(n):
                    if not cut[e]:
                        continue
                    b = -1
                    for i, bn in enumerate(bs.bins):
                        ok = True
                        for v, (lo, hi) in bn.edges.items():
                            x = col(v)[e]
                            ok &= (x == lo) if lo == hi else (lo <= x < hi)
                        if ok:
                            b = i
                            break
                    if b < 0:
                        continue
                    base.append(st.weight[e] * wf[e])
                    bins.append(s.bin_offset + b)
                    c = 0
                    for d in prop.dial_sets:
                        if d.datasets is not None and ds["name"] not in d.datasets:
                            continue
                        if not select(d.condition, col, n)[e]:
                            continue
                        if d.kind in ("leaf", "surface"):
                            t, beg, data = st.dial_leaf(d.leaf)
                            if beg[e + 1] == beg[e]:
                                continue
                            payload.append(np.asarray(data[beg[e]:beg[e + 1]]))
                        else:
                            t = N.DIAL_NORM
                            payload.append(np.zeros(0))
                        types.append(t)
                        inputs.append(d.input_index)
                        inputs2.append(d.input2_index)
                        c += 1
                    counts.append(c)
    return base, bins, types, inputs, payload, counts, inputs2


def test_config_assembly_matches_event_by_event_restatement(tmp_path):
    path = caller_fixture.write(str(tmp_path), n_mc=700, n_data=300)
    with open(path) as f:
        cfg = yaml.safe_load(f)
    prop = Propagator(cfg, str(tmp_path), build=False, binner=pyoracle.bin_events)
    t = prop.tables
    assert [p.name for p in prop.parameters] == ["norm_A", "norm_B", "spline_C", "spline_D",
                                                 "graph_E", "flux", "off", "Var_C", "Var_D"]
    np.testing.assert_array_equal(prop.prior_values(), [1, 1, 0, 0, 0, 1, 0, 0, 0])
    assert [p.free for p in prop.parameters] == [False] * 7 + [True, True]
    assert [s.name for s in prop.samples] == ["AB", "AB_negC"] and prop.n_bins == 72
    assert prop.reweight_cap == 3.5 and prop.llh_type == "LeastSquares"
    base, bins, types, inputs, payload, counts, inputs2 = _expected_tables(prop)
    np.testing.assert_array_equal(t.base_weight, base)
    np.testing.assert_array_equal(t.bin, bins)
    np.testing.assert_array_equal(np.diff(t.dial_begin), counts)
    np.testing.assert_array_equal(t.dial_type, types)
    np.testing.assert_array_equal(t.dial_input, inputs)
    np.testing.assert_array_equal(t.dial_input2, inputs2)
    assert {N.DIAL_BICUBIC, N.DIAL_BILINEAR} <= set(np.unique(t.dial_type).tolist())
    np.testing.assert_array_equal(np.diff(t.data_begin), [len(p) for p in payload])
    np.testing.assert_array_equal(t.dial_data, np.concatenate(payload))
    # one input per enabled dial set, with the configured supervisor / mirror
    inp = t.inputs
    # two surfaces on (Var_C, Var_D): an x and a y input each
    assert list(inp["parameter"]) == [0, 1, 2, 3, 4, 5, 7, 8, 7, 8]
    assert inp["min_response"][0] == 0.0 and inp["max_response"][2] == 2.5
    assert inp["flags"][3] == N.INPUT_NO_EXTRAPOLATION and inp["flags"][2] == 0
    assert inp["mirror_lower"][4] == -1.5 and inp["mirror_range"][4] == 3.0
    assert inp["min_response"][4] == 0.2 and np.isnan(inp["mirror_lower"][0])
    # penalty sets: the covariance set and the two bounds-only sets
    sets = prop.penalty_sets()
    assert len(sets) == 3
    with np.load(os.path.join(str(tmp_path), "cov.npz")) as z:
        cov = z["CovarianceInputCovariance"]
    np.testing.assert_allclose(sets[0]["inv_cov"] @ cov, np.eye(4), atol=1e-10)
    assert sets[0]["parameters"] == [0, 1, 2, 3] and sets[0]["prior"] == [1, 1, 0, 0]
    assert sets[1]["lower"][1] == 0.0 and sets[2]["lower"][0] == -4.0
    assert np.isnan(sets[2]["upper"][0])


def test_config_errors(tmp_path):
    path = caller_fixture.write(str(tmp_path), n_mc=50, n_data=10)
    with open(path) as f:
        cfg = yaml.safe_load(f)
    pc = cfg["fitterEngineConfig"]["propagatorConfig"]
    bad = yaml.safe_load(yaml.safe_dump(cfg))
    bad["fitterEngineConfig"]["likelihoodInterfaceConfig"]["jointProbabilityConfig"]["type"] = "Nope"
    with pytest.raises(ConfigError):
        Propagator(bad, str(tmp_path), build=False, binner=pyoracle.bin_events)
    bad = yaml.safe_load(yaml.safe_dump(cfg))
    bad["fitterEngineConfig"]["propagatorConfig"]["parameterSetListConfig"][1][
        "parameterDefinitions"][0]["dialSetDefinitions"][0]["dialsType"] = "Wiggle"
    with pytest.raises(ConfigError):
        Propagator(bad, str(tmp_path), build=False, binner=pyoracle.bin_events)
    assert pc["dataSetList"][0]["name"] == "TestSample"


def test_formula_not_binds_like_tformula():
    """TFormula's '!' is a tight unary operator: '!a == b' is '(!a) == b' and
    '!x > 0' is '(!x) > 0' (Python's 'not' would bind looser than both)."""
    cols = {"a": np.array([0.0, 1.0, 2.0, 0.0]), "b": np.array([1.0, 0.0, 0.0, 0.0]),
            "x": np.array([-1.0, 0.0, 3.0, 1.0])}
    col = cols.__getitem__
    np.testing.assert_array_equal(evaluate("!x > 0", col, 4), [0.0, 1.0, 0.0, 0.0])
    np.testing.assert_array_equal(evaluate("!(x > 0)", col, 4), [1.0, 1.0, 0.0, 0.0])
    np.testing.assert_array_equal(evaluate("![a] == [b]", col, 4), [1.0, 1.0, 1.0, 0.0])
    np.testing.assert_array_equal(evaluate("![a] == 0", col, 4), [0.0, 1.0, 1.0, 0.0])
    np.testing.assert_array_equal(evaluate("!a * 3", col, 4), [3.0, 0.0, 0.0, 3.0])
    np.testing.assert_array_equal(evaluate("!!a", col, 4), [0.0, 1.0, 1.0, 0.0])
    np.testing.assert_array_equal(evaluate("a != b", col, 4), [1.0, 1.0, 1.0, 0.0])
    with pytest.raises(FormulaError):
        evaluate("~a", col, 4)


def _first_dial_set(cfg):
    return cfg["fitterEngineConfig"]["propagatorConfig"]["parameterSetListConfig"][1][
        "parameterDefinitions"][0]["dialSetDefinitions"][0]


def test_min_spline_response_alias(tmp_path):
    """minimumSplineResponse is read as minDialResponse (DialCollection.cpp:42)."""
    path = caller_fixture.write(str(tmp_path), n_mc=50, n_data=10)
    with open(path) as f:
        cfg = yaml.safe_load(f)
    ds = _first_dial_set(cfg)
    del ds["minDialResponse"]
    ds["minimumSplineResponse"] = 0.35
    prop = Propagator(cfg, str(tmp_path), build=False, binner=pyoracle.bin_events)
    assert prop.tables.inputs["min_response"][4] == 0.35


@pytest.mark.parametrize("key,value", [
    ("parametersBinningPath", "bins.txt"), ("binning", "bins.txt"),
    ("binningFilePath", "bins.txt"), ("definitionRange", [-1.0, 1.0]),
    ("mirrorDefinitionRange", [-1.0, 1.0]), ("dialsFilePath", "d.root")])
def test_unsupported_dial_keys_are_rejected(tmp_path, key, value):
    """Binned dials, definition ranges and dial files would change which
    events a dial weights; a configuration using them is refused."""
    path = caller_fixture.write(str(tmp_path), n_mc=50, n_data=10)
    with open(path) as f:
        cfg = yaml.safe_load(f)
    _first_dial_set(cfg)[key] = value
    with pytest.raises(ConfigError, match=key):
        Propagator(cfg, str(tmp_path), build=False, binner=pyoracle.bin_events)

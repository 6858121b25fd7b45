"""YAML-driven caller of the cache: the parts of the reference's Propagator
and LikelihoodInterface that feed Cache::Manager and read its results.

    load_config(path) -> Propagator          (gundamFitter's config reading:
                                              fitterEngineConfig/propagatorConfig)
    Propagator.apply_parameters() -> Future  (Propagator::applyParameters,
                                              Propagator.cpp:117-155)
    Propagator.event_weights()               (CopyEventWeights into the events,
                                              FitterEngine.cpp:429-444)
    LikelihoodInterface.eval_likelihood()    (LikelihoodInterface.cpp:165-213:
                                              stat + per-set covariance penalty)

Inputs are event stores (gundam_amd/eventstore.py) in place of ROOT trees,
binning text files in the reference's format (gundam_amd/binning.py), and
covariance matrices as .npz arrays in place of ROOT TMatrixD objects.  The
configuration keys are the reference's; what is read:

    propagatorConfig:
      globalEventReweightCap: <number> | {isEnabled, value}
      dataSetList: [{name, isEnabled, selectedDataEntry,
                     mc: {filePathList, selectionCutFormula, nominalWeightFormula},
                     data: [{name, filePathList, selectionCutFormula,
                             nominalWeightFormula}]}]
      sampleSetConfig | fitSampleSetConfig:
        sampleList | fitSampleList: [{name, isEnabled, binning, selectionCutStr,
                                      datasets | dataSets}]
        llhStatFunction, llhConfig          (relocated keys, still read)
      parametersManagerConfig: {parameterSetList: [...]} | parameterSetListConfig: [...]
        - name, isEnabled, covarianceMatrixFilePath (.npz),
          covarianceMatrixTMatrixD, parameterPriorTVectorD,
          parameterNameTObjArray (array names inside the .npz),
          parameterDefinitions: [{parameterName | parameterIndex, isEnabled,
              isFixed, isFree, priorType (Flat: free), isPenaltyDisabled,
              priorValue, parameterLimits,
              dialSetDefinitions: [{dialType | dialsType | type, isEnabled,
                  dialLeafName | treeExpression, applyOnDataSets,
                  applyCondition, minDialResponse, maxDialResponse,
                  useMirrorDial, mirrorLowEdge, mirrorHighEdge,
                  allowDialExtrapolation}]}],
          dialSetDefinitions: [{... as above, dialInputList: [{name}, ...],
                                dialSubType}]   (set level: one input, or two
                                                 for dialType Surface)
    likelihoodInterfaceConfig:
      jointProbabilityConfig: {type, <llhConfig keys>}
      applyInfinitePenaltyOnOutOfBoundPar

``${NAME}`` in strings is replaced from the environment; CONFIG_DIR defaults
to the configuration file's directory.  Dial types: Normalization (no leaf),
Spline / Graph / Shift (the payload and its type come from the store leaf),
Surface (a bilinear or bicubic store leaf, two inputs).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import yaml

from . import _native as N
from .binning import BinSet, read_binning_text
from .cache_manager import CacheManager, Future
from .eventstore import EventStore
from .formula import evaluate, select
from .tables import INPUT_DTYPE, Tables

# llhStatFunction / jointProbabilityConfig.type -> cache likelihood kind
# (JointProbability.cpp:61-94)
JOINT_PROBABILITIES = {
    "PoissonLLH": "poisson", "BarlowLLH": "barlow_classic",
    "BarlowLLH_BANFF_OA2020": "barlow_oa2020", "BarlowLLH_BANFF_OA2021": "barlow",
    "BarlowLLH_BANFF_OA2021_SFGD": "barlow_oa2021_sfgd", "Chi2": "chi2",
    "LeastSquares": "least_squares",
}


class ConfigError(ValueError):
    pass


def _expand(obj, env):
    if isinstance(obj, str):
        return re.sub(r"\$\{([A-Za-z_][A-Za-z0-9_]*)\}",
                      lambda m: env.get(m.group(1), m.group(0)), obj)
    if isinstance(obj, list):
        return [_expand(x, env) for x in obj]
    if isinstance(obj, dict):
        return {k: _expand(v, env) for k, v in obj.items()}
    return obj


def _get(d: dict, *names, default=None):
    for n in names:
        if isinstance(d, dict) and n in d:
            return d[n]
    return default


@dataclass
class DialSet:
    kind: str                    # "normalization", "leaf" or "surface" (2-D leaf)
    leaf: Optional[str]
    datasets: Optional[List[str]]  # None: every dataset
    condition: Optional[str]
    input_index: int
    input2_index: int = -1       # the y input of a surface


@dataclass
class Parameter:
    name: str
    index: int                   # position in the parameter vector
    prior: float
    lower: float = np.nan
    upper: float = np.nan
    enabled: bool = True
    fixed: bool = False
    free: bool = False
    penalty_disabled: bool = False
    dial_sets: List[DialSet] = field(default_factory=list)


@dataclass
class ParameterSet:
    name: str
    parameters: List[Parameter]
    covariance: Optional[np.ndarray] = None  # over `parameters`, full

    def correlated(self) -> List[int]:
        """Positions of the valid correlated parameters
        (ParameterSet::isValidCorrelatedParameter)."""
        return [i for i, p in enumerate(self.parameters)
                if p.enabled and not p.penalty_disabled and not p.free]


@dataclass
class Sample:
    name: str
    binning: BinSet
    bin_offset: int
    selection: Optional[str]
    datasets: Optional[List[str]]


class Propagator:
    """Parameters, samples and the cache built from a configuration."""

    def __init__(self, cfg: dict, base_dir: str, device: int = 0, max_points: int = 1,
                 precision: str = "fp64", build: bool = True, binner=None):
        """build=False reads the configuration and assembles the tables
        without a device cache; binner(values [n][v], edges [b][v][2]) -> bins
        replaces the device binning (gcm_bin_events) for that case."""
        self._binner = binner
        env = dict(os.environ)
        env.setdefault("CONFIG_DIR", base_dir)
        cfg = _expand(cfg, env)
        fe = _get(cfg, "fitterEngineConfig", default={}) or {}
        pc = _get(fe, "propagatorConfig", default=None) or _get(cfg, "propagatorConfig")
        if pc is None:
            raise ConfigError("no fitterEngineConfig/propagatorConfig")
        lic = _get(fe, "likelihoodInterfaceConfig", default={}) or \
            _get(cfg, "likelihoodInterfaceConfig", default={}) or {}
        self.base_dir = base_dir
        self.device = device
        ssc = _get(pc, "sampleSetConfig", "fitSampleSetConfig", default={}) or {}
        self._read_datasets(_get(pc, "dataSetList", default=None) or
                            _get(ssc, "dataSetList", default=None) or
                            _get(lic, "dataSetList", default=[]))
        self._read_samples(_get(ssc, "sampleList", "fitSampleList", default=[]))
        pm = _get(pc, "parametersManagerConfig", default={}) or {}
        self._read_parameters(_get(pm, "parameterSetList", default=None) or
                              _get(pc, "parameterSetListConfig", default=[]))
        cap = _get(pc, "globalEventReweightCap", default=None)
        if isinstance(cap, dict):
            cap = cap.get("value") if cap.get("isEnabled", True) else None
        self.reweight_cap = float(cap) if cap is not None else np.inf
        jp = dict(_get(lic, "jointProbabilityConfig", default={}) or {})
        jp.update(_get(ssc, "llhConfig", default={}) or {})
        jtype = jp.get("type") or _get(ssc, "llhStatFunction", default=None) or \
            _get(pc, "llhStatFunction", default="PoissonLLH")
        if jtype not in JOINT_PROBABILITIES:
            raise ConfigError(f"joint probability {jtype} not supported "
                              f"(have {sorted(JOINT_PROBABILITIES)})")
        self.llh_type, self.llh_options = jtype, jp
        if _get(ssc, "dataEventType", default=None) == "Asimov":
            for ds in self.datasets:
                ds["selected_data"] = "Asimov"
        self.infinite_out_of_bounds = bool(_get(lic, "applyInfinitePenaltyOnOutOfBoundPar",
                                                default=False))
        self._assemble()
        self.values = self.prior_values()
        self.cache = None
        if build:
            self.cache = CacheManager(device, precision=precision)
            self.cache.build(self.tables, max_points=max_points)
            self.cache.set_event_weight_copy(True)

    # -- configuration ------------------------------------------------------
    def _path(self, p: str) -> str:
        return p if os.path.isabs(p) else os.path.join(self.base_dir, p)

    def _read_datasets(self, lst):
        self.datasets = []
        for d in lst or []:
            if not d.get("isEnabled", True):
                continue
            mc = d.get("mc", {}) or {}
            self.datasets.append({
                "name": d["name"],
                "mc": [EventStore(self._path(p)) for p in mc.get("filePathList", [])],
                "mc_cut": mc.get("selectionCutFormula"),
                "mc_weight": mc.get("nominalWeightFormula", mc.get("nominalTreeWeightFormula")),
                "selected_data": d.get("selectedDataEntry", "Asimov"),
                "data": {e["name"]: e for e in (d.get("data") or [])},
            })
        if not self.datasets:
            raise ConfigError("no enabled dataset")

    def _read_samples(self, lst):
        self.samples = []
        off = 0
        for s in lst or []:
            if not s.get("isEnabled", True):
                continue
            b = _get(s, "binning", "binningFile", "binningFilePath")
            if not isinstance(b, str):
                raise ConfigError(f"sample {s.get('name')}: binning must be a text file path")
            bs = read_binning_text(self._path(b))
            self.samples.append(Sample(s["name"], bs, off, _get(s, "selectionCutStr", "selectionCuts"),
                                       _get(s, "datasets", "dataSets")))
            off += bs.n_bins
        if not self.samples:
            raise ConfigError("no enabled sample")
        self.n_bins = off

    def _read_parameters(self, lst):
        self.parameter_sets: List[ParameterSet] = []
        self.parameters: List[Parameter] = []
        self._inputs = []
        self.dial_sets: List[DialSet] = []  # collection order
        for ps in lst or []:
            if not ps.get("isEnabled", True):
                continue
            names, priors, cov = None, None, None
            if ps.get("covarianceMatrixFilePath"):
                with np.load(self._path(ps["covarianceMatrixFilePath"]), allow_pickle=False) as z:
                    cov = np.array(z[ps.get("covarianceMatrixTMatrixD", "covariance")], dtype=np.float64)
                    if ps.get("parameterPriorTVectorD"):
                        priors = np.array(z[ps["parameterPriorTVectorD"]], dtype=np.float64)
                    if ps.get("parameterNameTObjArray"):
                        names = [str(x) for x in z[ps["parameterNameTObjArray"]]]
                if cov.ndim != 2 or cov.shape[0] != cov.shape[1]:
                    raise ConfigError(f"parameter set {ps['name']}: covariance is not square")
                if not np.allclose(cov, cov.T, rtol=0, atol=0):
                    raise ConfigError(f"parameter set {ps['name']}: covariance is not symmetric")
            defs = ps.get("parameterDefinitions") or []
            if cov is None:
                # no covariance: the definitions are the parameters, in order
                pars = [Parameter(name=dd.get("parameterName", f"{ps['name']}_{i}"), index=-1,
                                  prior=0.0) for i, dd in enumerate(defs)]
                pairs = list(zip(pars, defs))
            else:
                pars = [Parameter(name=(names[i] if names else f"{ps['name']}_{i}"), index=-1,
                                  prior=float(priors[i]) if priors is not None else 0.0)
                        for i in range(cov.shape[0])]
                pairs = []
                for i, dd in enumerate(defs):
                    if "parameterIndex" in dd:
                        j = int(dd["parameterIndex"])
                    elif names and "parameterName" in dd:
                        if dd["parameterName"] not in names:
                            raise ConfigError(f"{ps['name']}: unknown parameter {dd['parameterName']}")
                        j = names.index(dd["parameterName"])
                    else:
                        j = i
                    if not 0 <= j < len(pars):
                        raise ConfigError(f"{ps['name']}: parameter definition {i} out of range")
                    if "parameterName" in dd:
                        pars[j].name = dd["parameterName"]
                    pairs.append((pars[j], dd))
            dial_defs = {}
            for p, dd in pairs:
                p.enabled = bool(dd.get("isEnabled", True))
                p.fixed = bool(dd.get("isFixed", False))
                # a flat prior makes the parameter free (Parameter.cpp:64)
                p.free = bool(dd.get("isFree", False)) or dd.get("priorType") == "Flat"
                p.penalty_disabled = bool(dd.get("isPenaltyDisabled", False))
                if "priorValue" in dd:
                    p.prior = float(dd["priorValue"])
                lim = dd.get("parameterLimits")
                if lim is not None:
                    p.lower = np.nan if lim[0] is None else float(lim[0])
                    p.upper = np.nan if lim[1] is None else float(lim[1])
                dial_defs[id(p)] = (dd.get("dialSetDefinitions") or []) if p.enabled else []
            for p in pars:
                p.index = len(self.parameters)
                self.parameters.append(p)
                for dd in dial_defs.get(id(p), []):
                    if not dd.get("isEnabled", True):
                        continue
                    self._add_dial_set(p, dd)
            # set-level dial sets name their inputs (DialCollection.cpp:78-100)
            for dd in ps.get("dialSetDefinitions") or []:
                if not dd.get("isEnabled", True):
                    continue
                names_in = [x.get("name") for x in (dd.get("dialInputList") or [])]
                ins = []
                for nm in names_in:
                    hit = [q for q in pars if q.name == nm]
                    if not hit:
                        raise ConfigError(f"{ps['name']}: dialInputList names unknown parameter {nm}")
                    ins.append(hit[0])
                if len(ins) == 1:
                    if ins[0].enabled:
                        self._add_dial_set(ins[0], dd)
                elif len(ins) == 2:
                    if ins[0].enabled and ins[1].enabled:
                        self._add_dial_set(ins[0], dd, ins[1])
                else:
                    raise ConfigError(f"{ps['name']}: a dial set needs one or two dialInputList entries")
            self.parameter_sets.append(ParameterSet(ps["name"], pars, cov))
        if not self.parameters:
            raise ConfigError("no enabled parameter")

    # dial-set keys of the reference (DialCollection.cpp:30-60) that change
    # which events a dial applies to or what it returns, and are not built
    # here: a configuration that uses them is rejected rather than loaded with
    # different weights
    _UNSUPPORTED_DIAL_KEYS = ("parametersBinningPath", "binning", "binningFilePath",
                              "definitionRange", "mirrorDefinitionRange", "dialsFilePath",
                              "dialsList", "dialsTreePath", "dialsDefinitions", "tableConfig")

    def _add_dial_set(self, p: Parameter, dd: dict, p2: Optional[Parameter] = None):
        bad = [k for k in self._UNSUPPORTED_DIAL_KEYS if k in dd]
        if bad:
            raise ConfigError(f"parameter {p.name}: dial-set key(s) {', '.join(bad)} "
                              f"not supported (binned dials, definition ranges and "
                              f"dial files are not built)")
        t = str(_get(dd, "dialType", "type", "dialsType", default="")).lower()
        leaf = _get(dd, "dialLeafName", "treeExpression")
        if (t == "surface") != (p2 is not None):
            raise ConfigError(f"parameter {p.name}: Surface dials take exactly two inputs")
        if t in ("normalization", "norm"):
            kind = "normalization"
        elif t == "surface":
            if not leaf:
                raise ConfigError(f"parameter {p.name}: Surface dials need dialLeafName")
            sub = str(_get(dd, "dialSubType", "options", default="")).lower()
            if sub and sub not in ("bilinear", "bicubic"):
                raise ConfigError(f"parameter {p.name}: surface sub-type {sub} not supported")
            kind = "surface"
        elif t in ("tabulated", "kriged"):
            raise ConfigError(f"parameter {p.name}: dial type {t} is not read from "
                              f"configuration files (use the C ABI directly)")
        elif t in ("spline", "graph", "shift"):
            if not leaf:
                raise ConfigError(f"parameter {p.name}: dial type {t} needs dialLeafName")
            kind = "leaf"
        else:
            raise ConfigError(f"parameter {p.name}: unknown dial type {t!r}")
        nan = np.nan
        mir_lo, mir_r = nan, nan
        if dd.get("useMirrorDial", False):
            lo, hi = float(dd["mirrorLowEdge"]), float(dd["mirrorHighEdge"])
            if hi < lo:
                raise ConfigError(f"parameter {p.name}: mirrorHighEdge < mirrorLowEdge")
            mir_lo, mir_r = lo, hi - lo
        flags = 0 if dd.get("allowDialExtrapolation", True) else N.INPUT_NO_EXTRAPOLATION
        # DialCollection.cpp:453,915-917: minDialResponse defaults to 0;
        # minimumSplineResponse is its alias (DialCollection.cpp:42)
        mn = _get(dd, "minDialResponse", "minimumSplineResponse", default=0.0)
        mx = dd.get("maxDialResponse", None)
        self._inputs.append((p.index, flags, mir_lo, mir_r,
                             nan if mn is None else float(mn), nan if mx is None else float(mx)))
        in1 = len(self._inputs) - 1
        in2 = -1
        if p2 is not None:
            # the y input: the same mirror edges (they apply to every input of
            # the collection); the response clamp is the x input's
            self._inputs.append((p2.index, flags, mir_lo, mir_r, nan, nan))
            in2 = len(self._inputs) - 1
        ds = dd.get("applyOnDataSets")
        if ds is not None and ("*" in ds or "all" in ds):
            ds = None
        dset = DialSet(kind, leaf, ds, _get(dd, "applyCondition", "applyConditions"), in1, in2)
        p.dial_sets.append(dset)
        self.dial_sets.append(dset)

    # -- event and dial tables ---------------------------------------------
    def _bin(self, sample: Sample, store: EventStore, mask: np.ndarray) -> np.ndarray:
        vars_ = sample.binning.variables()
        idx = np.flatnonzero(mask)
        out = np.full(store.n_events, -1, np.int32)
        if idx.size == 0 or sample.binning.n_bins == 0:
            return out
        vals = np.ascontiguousarray(np.stack([np.asarray(store.column(v))[idx] for v in vars_], 1)
                                    if vars_ else np.zeros((idx.size, 0)))
        edges = np.ascontiguousarray(sample.binning.edge_array(vars_))
        if self._binner is not None:
            out[idx] = self._binner(vals, edges)
            return out
        b = np.empty(idx.size, np.int32)
        N.check(N.lib().gcm_bin_events(self.device, idx.size, len(vars_), vals.ctypes.data,
                                       sample.binning.n_bins, edges.ctypes.data, b.ctypes.data))
        out[idx] = b
        return out

    def _assemble(self):
        """One cache event per (sample, dataset, file, event) that passes the
        dataset and sample selections and falls in a bin of the sample; its
        dials in collection order (parameter sets, parameters, dial sets)."""
        blocks = []            # per (sample, dataset, file) block: arrays
        self.event_index = []  # per sample: (dataset, file, event indices)
        for s in self.samples:
            rows = []
            for di, ds in enumerate(self.datasets):
                if s.datasets is not None and ds["name"] not in s.datasets:
                    continue
                for fi, st in enumerate(ds["mc"]):
                    blk = self._block(s, ds, st)
                    if blk is None:
                        continue
                    blocks.append(blk)
                    rows.append((di, fi, blk["keep"]))
            self.event_index.append(rows)
        # concatenate the blocks
        cat = lambda k, dt: (np.concatenate([b[k] for b in blocks]).astype(dt)
                             if blocks else np.zeros(0, dt))
        n_dials_per_event = cat("n_dials", np.int64)
        lens = cat("lens", np.int64)
        inputs = np.zeros(len(self._inputs), dtype=INPUT_DTYPE)
        for i, (par, fl, mlo, mr, mn, mx) in enumerate(self._inputs):
            inputs[i] = (par, fl, mlo, mr, mn, mx)
        self.tables = Tables(
            base_weight=cat("base", np.float64),
            bin=cat("bin", np.int32),
            dial_begin=np.concatenate([[0], np.cumsum(n_dials_per_event)]).astype(np.int64),
            dial_type=cat("dtype", np.uint8),
            dial_input=cat("dinput", np.int32),
            data_begin=np.concatenate([[0], np.cumsum(lens)]).astype(np.int64),
            dial_data=cat("payload", np.float64),
            inputs=inputs, table_begin=np.zeros(1, np.int64),
            n_parameters=len(self.parameters), n_bins=self.n_bins,
            reweight_cap=self.reweight_cap, dial_input2=cat("dinput2", np.int32))
        self.n_events = int(self.tables.n_events)

    def _block(self, s: Sample, ds: dict, st: EventStore):
        n = st.n_events
        m = select(ds["mc_cut"], st.column, n) & select(s.selection, st.column, n)
        b = self._bin(s, st, m)
        keep = np.flatnonzero(b >= 0)
        if keep.size == 0:
            return None
        w = np.array(st.weight, dtype=np.float64)
        if ds["mc_weight"]:
            w = w * evaluate(ds["mc_weight"], st.column, n)
        # the dial sets that apply to this dataset, in collection order
        sets = [d for d in self.dial_sets if d.datasets is None or ds["name"] in d.datasets]
        on = np.zeros((keep.size, len(sets)), dtype=bool)
        types = np.zeros(len(sets), np.uint8)
        inp = np.zeros(len(sets), np.int32)
        inp2 = np.full(len(sets), -1, np.int32)
        leaves = []
        for j, d in enumerate(sets):
            m_j = select(d.condition, st.column, n)[keep]
            inp[j] = d.input_index
            inp2[j] = d.input2_index
            if d.kind in ("leaf", "surface"):
                t, beg, data = st.dial_leaf(d.leaf)
                is2d = t in (N.DIAL_BILINEAR, N.DIAL_BICUBIC)
                if is2d != (d.kind == "surface"):
                    raise ConfigError(f"dial leaf {d.leaf}: a {d.kind} dial set needs "
                                      f"{'a 2-D' if d.kind == 'surface' else 'a 1-D'} leaf")
                beg = np.asarray(beg)
                m_j &= (beg[keep + 1] - beg[keep]) > 0
                types[j] = t
                leaves.append((beg, data))
            else:
                types[j] = N.DIAL_NORM
                leaves.append(None)
            on[:, j] = m_j
        ev, sj = np.nonzero(on)  # event-major, collection order inside an event
        lens = np.zeros(ev.size, np.int64)
        starts = np.zeros(ev.size, np.int64)
        for j, lf in enumerate(leaves):
            if lf is None:
                continue
            sel = sj == j
            e = keep[ev[sel]]
            starts[sel] = lf[0][e]
            lens[sel] = lf[0][e + 1] - lf[0][e]
        off = np.concatenate([[0], np.cumsum(lens)])
        payload = np.empty(int(off[-1]))
        for j, lf in enumerate(leaves):
            if lf is None:
                continue
            sel = np.flatnonzero(sj == j)
            if sel.size == 0:
                continue
            # segment i of this leaf: data[starts_i : starts_i + L_i] goes to
            # payload[off_i : off_i + L_i]; r runs over the concatenated segments
            L = lens[sel]
            cx = np.concatenate([[0], np.cumsum(L)[:-1]])
            r = np.arange(int(L.sum()))
            payload[np.repeat(off[sel] - cx, L) + r] = \
                np.asarray(lf[1])[np.repeat(starts[sel] - cx, L) + r]
        return {"keep": keep, "base": w[keep], "bin": s.bin_offset + b[keep],
                "n_dials": np.bincount(ev, minlength=keep.size), "dtype": types[sj],
                "dinput": inp[sj], "dinput2": inp2[sj], "lens": lens, "payload": payload}

    # -- parameters and evaluation ------------------------------------------
    def prior_values(self) -> np.ndarray:
        return np.array([p.prior for p in self.parameters], dtype=np.float64)

    def set_parameter_values(self, values) -> None:
        v = np.asarray(values, dtype=np.float64)
        if v.shape[-1] != len(self.parameters):
            raise ValueError(f"{v.shape[-1]} values for {len(self.parameters)} parameters")
        self.values = v.copy()

    def apply_parameters(self) -> Future:
        """Propagator::applyParameters: reweight and refill every sample
        histogram for the current parameter values (deferred)."""
        self.cache.set_parameters(self.values)
        return self.cache.fill()

    def sample_histogram(self, sample: int | str, k: int = 0):
        """(sum w, sqrt(sum w^2)) of one sample's bins."""
        s = self._sample(sample)
        sw, e = self.cache.copy_histogram_contents(k)
        sl = slice(s.bin_offset, s.bin_offset + s.binning.n_bins)
        return sw[sl].copy(), e[sl].copy()

    def _sample(self, sample) -> Sample:
        if isinstance(sample, str):
            return next(s for s in self.samples if s.name == sample)
        return self.samples[sample]

    def event_weights(self) -> Dict[str, list]:
        """Per sample: [(dataset name, file index, event indices, weights)],
        the current weight of every cache event (CopyEventWeights)."""
        w = self.cache.copy_event_weights(np.full(self.tables.n_events, np.nan))
        out, pos = {}, 0
        for s, rows in zip(self.samples, self.event_index):
            lst = []
            for di, fi, keep in rows:
                lst.append((self.datasets[di]["name"], fi, keep, w[pos:pos + keep.size].copy()))
                pos += keep.size
            out[s.name] = lst
        return out

    def penalty_sets(self) -> List[dict]:
        """gcm_penalty_set descriptions: per parameter set with a covariance,
        its valid correlated parameters with the inverse of their stripped
        covariance (ParameterSet.cpp:225-247); with
        applyInfinitePenaltyOnOutOfBoundPar, a bounds-only set over all of its
        parameters (ParameterSet::hasOutOfBoundsParameters)."""
        sets = []
        for ps in self.parameter_sets:
            par = ps.parameters
            corr = ps.correlated()
            if ps.covariance is not None and corr:
                c = ps.covariance[np.ix_(corr, corr)]
                sets.append({"parameters": [par[i].index for i in corr],
                             "prior": [par[i].prior for i in corr],
                             "inv_cov": np.linalg.inv(c)})
            if self.infinite_out_of_bounds:
                sets.append({"parameters": [p.index for p in par],
                             "prior": [p.prior for p in par],
                             "lower": [p.lower for p in par],
                             "upper": [p.upper for p in par]})
        return sets

    def data_histogram(self) -> np.ndarray:
        """Data per bin: the Asimov prediction at the prior for datasets whose
        selectedDataEntry is Asimov (or empty), else the selected data
        entry's events binned with their weights
        (LikelihoodInterface.cpp:523-535)."""
        is_asimov = [ds["selected_data"] in (None, "", "Asimov") for ds in self.datasets]
        asimov = None
        if any(is_asimov):
            keep_vals = self.values
            self.set_parameter_values(self.prior_values())
            self.apply_parameters().get()
            if all(is_asimov):  # the prediction at the prior, as the device summed it
                self.set_parameter_values(keep_vals)
                return self.cache.hist(0)[0].copy()
            asimov = self._per_dataset_hist()
            self.set_parameter_values(keep_vals)
        data = np.zeros(self.n_bins)
        for di, ds in enumerate(self.datasets):
            sel = ds["selected_data"]
            if sel in (None, "", "Asimov"):
                data += asimov[di]
                continue
            if sel not in ds["data"]:
                raise ConfigError(f"dataset {ds['name']}: no data entry {sel}")
            ent = ds["data"][sel]
            for p in ent.get("filePathList", []):
                st = EventStore(self._path(p))
                w = np.array(st.weight, dtype=np.float64)
                if ent.get("nominalWeightFormula"):
                    w = w * evaluate(ent["nominalWeightFormula"], st.column, st.n_events)
                for s in self.samples:
                    if s.datasets is not None and ds["name"] not in s.datasets:
                        continue
                    m = select(ent.get("selectionCutFormula"), st.column, st.n_events) & \
                        select(s.selection, st.column, st.n_events)
                    b = self._bin(s, st, m)
                    ok = b >= 0
                    data += np.bincount(s.bin_offset + b[ok], weights=w[ok], minlength=self.n_bins)
        return data

    def _per_dataset_hist(self) -> List[np.ndarray]:
        """Weighted prediction per dataset (for Asimov data by dataset)."""
        ev = self.event_weights()
        out = [np.zeros(self.n_bins) for _ in self.datasets]
        names = [d["name"] for d in self.datasets]
        pos = 0
        for s, rows in zip(self.samples, self.event_index):
            for (dname, fi, keep, w) in ev[s.name]:
                b = self.tables.bin[pos:pos + keep.size]
                out[names.index(dname)] += np.bincount(b, weights=w, minlength=self.n_bins)
                pos += keep.size
        return out


class LikelihoodInterface:
    """Stat likelihood of every sample plus the prior penalty of every
    parameter set, both evaluated on the device (LikelihoodInterface.cpp:
    165-213).  ``buffer`` holds the last (stat, penalty, total)."""

    def __init__(self, prop: Propagator, data: Optional[np.ndarray] = None):
        self.prop = prop
        cm = prop.cache
        self.data = prop.data_histogram() if data is None else np.asarray(data, np.float64)
        opts = prop.llh_options
        kind = JOINT_PROBABILITIES[prop.llh_type]
        ignore = None
        if opts.get("ignoreBinsWithZeroPredictionAtPrior", False):
            prop.set_parameter_values(prop.prior_values())
            prop.apply_parameters().get()
            ignore = cm.zero_prediction_mask()
        nominal_err = None
        if opts.get("BBNoUpdateWeights", False):
            prop.set_parameter_values(prop.prior_values())
            prop.apply_parameters().get()
            nominal_err = cm.copy_histogram_contents()[1].copy()
        cm.set_likelihood(
            kind, self.data, None, None,
            use_poisson_likelihood=bool(opts.get("usePoissonLikelihood", False)),
            fractional_error_limit=float(opts.get("fractionalErrorLimit", 1e150)),
            expected_value_minimum=float(opts.get("expectedValueMinimum", -1.0)),
            lsq_poissonian_approximation=bool(opts.get("lsqPoissonianApproximation", False)),
            nominal_err=nominal_err, ignore_bins=ignore)
        self.throw_if_inf = bool(opts.get("throwIfInfLlh", kind.startswith("barlow")
                                          and kind != "barlow_classic"))
        self.penalty_sets = prop.penalty_sets()
        cm.set_penalty_sets(self.penalty_sets)
        self.buffer = {"stat": np.nan, "penalty": np.nan, "total": np.nan}
        prop.set_parameter_values(prop.prior_values())

    def eval_likelihood(self, future: Optional[Future] = None) -> float:
        if future is not None and future.valid():
            future.get()
        tot, st, pen = self.prop.cache.llh()
        if self.throw_if_inf and not np.isfinite(st[0]):
            raise FloatingPointError(f"non-finite stat likelihood {st[0]}")
        self.buffer = {"stat": float(st[0]), "penalty": float(pen[0]), "total": float(tot[0])}
        return self.buffer["total"]

    def propagate_and_eval_likelihood(self) -> float:
        return self.eval_likelihood(self.prop.apply_parameters())


def load_config(path: str, device: int = 0, max_points: int = 1, precision: str = "fp64"):
    """(Propagator, LikelihoodInterface) for a YAML configuration file."""
    with open(path) as f:
        cfg = yaml.safe_load(f)
    prop = Propagator(cfg, os.path.dirname(os.path.abspath(path)), device=device,
                      max_points=max_points, precision=precision)
    return prop, LikelihoodInterface(prop)

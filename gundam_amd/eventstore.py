"""On-disk event/dial store: the input format of the YAML-driven loader
(gundam_amd/config.py), in place of the reference's ROOT trees
(DataDispenser reads the events, their variables and per-event dial objects
from trees named in the dataset configuration).

A store is a directory of NumPy ``.npy`` files (memory-mapped on read, no
pickles) with a ``manifest.json``:

    manifest.json    {"format": "gcm-event-store", "version": 1,
                      "n_events": N, "variables": [names...],
                      "dials": {leaf: {"type": "<dial type>"}}}
    variables.npy    float64 [N][V]   event variables (selection, binning)
    weight.npy       float64 [N]      nominal event weight
    dial_<leaf>_begin.npy  int64 [N+1]  per-event payload ranges into
    dial_<leaf>_data.npy   float64      the payloads; an empty range means
                                        the event has no dial of that leaf

Dial payloads are the cache's record layouts (include/gcm.h: the reference's
DialBase::getDialData() layouts); ``type`` is one of DIAL_TYPE_NAMES.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Sequence

import numpy as np

from . import _native as N

FORMAT = "gcm-event-store"
VERSION = 1

# dial type names of the store (lower case) -> GCM_DIAL_*
DIAL_TYPE_NAMES = {
    "normalization": N.DIAL_NORM, "monotonic": N.DIAL_MONOTONIC,
    "uniform": N.DIAL_UNIFORM, "general": N.DIAL_GENERAL,
    "compact": N.DIAL_COMPACT, "graph": N.DIAL_GRAPH, "shift": N.DIAL_SHIFT,
    "bilinear": N.DIAL_BILINEAR, "bicubic": N.DIAL_BICUBIC,
}


class StoreError(ValueError):
    pass


class EventStore:
    """A read-only view of one store (arrays memory-mapped)."""

    def __init__(self, path: str):
        self.path = path
        mf = os.path.join(path, "manifest.json")
        try:
            with open(mf) as f:
                man = json.load(f)
        except FileNotFoundError:
            raise StoreError(f"{path}: no manifest.json") from None
        if man.get("format") != FORMAT or man.get("version") != VERSION:
            raise StoreError(f"{path}: not a {FORMAT} v{VERSION} store")
        self.manifest = man
        self.n_events = int(man["n_events"])
        self.variable_names = list(man["variables"])
        self.variables = self._load("variables.npy")
        self.weight = self._load("weight.npy")
        if self.variables.shape != (self.n_events, len(self.variable_names)):
            raise StoreError(f"{path}: variables.npy has shape {self.variables.shape}")
        if self.weight.shape != (self.n_events,):
            raise StoreError(f"{path}: weight.npy has shape {self.weight.shape}")
        self.dial_types: Dict[str, int] = {}
        for leaf, spec in man.get("dials", {}).items():
            t = spec.get("type", "").lower()
            if t not in DIAL_TYPE_NAMES:
                raise StoreError(f"{path}: dial leaf {leaf}: unknown type {spec.get('type')}")
            self.dial_types[leaf] = DIAL_TYPE_NAMES[t]

    def _load(self, name: str) -> np.ndarray:
        return np.load(os.path.join(self.path, name), mmap_mode="r", allow_pickle=False)

    def column(self, name: str) -> np.ndarray:
        try:
            return self.variables[:, self.variable_names.index(name)]
        except ValueError:
            raise StoreError(f"{self.path}: no variable {name}") from None

    def dial_leaf(self, leaf: str):
        """(type, begin [N+1], data) of one dial leaf."""
        if leaf not in self.dial_types:
            raise StoreError(f"{self.path}: no dial leaf {leaf}")
        begin = self._load(f"dial_{leaf}_begin.npy")
        data = self._load(f"dial_{leaf}_data.npy")
        if begin.shape != (self.n_events + 1,) or begin[0] != 0 or begin[-1] != data.shape[0]:
            raise StoreError(f"{self.path}: dial leaf {leaf}: inconsistent offsets")
        return self.dial_types[leaf], begin, data


def write_store(path: str, variables: Dict[str, Sequence[float]], weight: Sequence[float],
                dials: Optional[Dict[str, tuple]] = None) -> EventStore:
    """Write a store.  variables: name -> [N] values; dials: leaf ->
    (type name, payloads) with payloads a list of N arrays (empty: no dial)
    or a (begin [N+1], data) pair."""
    names = list(variables)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    n = w.shape[0]
    cols = [np.asarray(variables[k], dtype=np.float64) for k in names]
    for k, c in zip(names, cols):
        if c.shape != (n,):
            raise StoreError(f"variable {k}: {c.shape[0] if c.ndim else 0} values for {n} events")
    os.makedirs(path, exist_ok=True)
    np.save(os.path.join(path, "variables.npy"),
            np.stack(cols, axis=1) if cols else np.zeros((n, 0)))
    np.save(os.path.join(path, "weight.npy"), w)
    man_dials = {}
    for leaf, (tname, payloads) in (dials or {}).items():
        if tname.lower() not in DIAL_TYPE_NAMES:
            raise StoreError(f"dial leaf {leaf}: unknown type {tname}")
        if isinstance(payloads, tuple):
            begin, data = (np.asarray(payloads[0], np.int64), np.asarray(payloads[1], np.float64))
        else:
            if len(payloads) != n:
                raise StoreError(f"dial leaf {leaf}: {len(payloads)} payloads for {n} events")
            lens = np.array([len(p) for p in payloads], np.int64)
            begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            data = (np.concatenate([np.asarray(p, np.float64) for p in payloads])
                    if n and lens.sum() else np.zeros(0))
        np.save(os.path.join(path, f"dial_{leaf}_begin.npy"), begin)
        np.save(os.path.join(path, f"dial_{leaf}_data.npy"), data)
        man_dials[leaf] = {"type": tname.lower()}
    with open(os.path.join(path, "manifest.json"), "w") as f:
        json.dump({"format": FORMAT, "version": VERSION, "n_events": int(n),
                   "variables": names, "dials": man_dials}, f, indent=1)
    return EventStore(path)

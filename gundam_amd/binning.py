"""Sample binnings in the reference's text format (BinSet::readTxtBinningDefinition,
src/Utils/src/BinSet.cpp:143-257), as the edge arrays gcm_bin_events takes.

Format: '#' starts a comment; blank lines are skipped; a line
``variables: A A B`` declares the columns of the following bin lines (a name
given twice is a [min, max) range, once a point value min == max); every
other line is one bin.  ``variables:`` may appear again, and the bins after it
use the new variables.  Each bin's edges are sorted by variable name,
case-insensitively (BinSet::sortBinEdges, BinSet.cpp:115-133), which does not
change which events a bin holds.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np


@dataclass
class Bin:
    edges: Dict[str, Tuple[float, float]]  # variable -> (min, max); min == max: point


@dataclass
class BinSet:
    bins: List[Bin] = field(default_factory=list)

    @property
    def n_bins(self) -> int:
        return len(self.bins)

    def variables(self) -> List[str]:
        """Every variable some bin constrains, in order of first appearance."""
        out: List[str] = []
        for b in self.bins:
            for v in b.edges:
                if v not in out:
                    out.append(v)
        return out

    def edge_array(self, variables: List[str]) -> np.ndarray:
        """[n_bins][n_vars][2] for gcm_bin_events; a variable a bin does not
        constrain gets (-inf, +inf), which every finite value passes."""
        out = np.empty((self.n_bins, len(variables), 2))
        out[..., 0] = -np.inf
        out[..., 1] = np.inf
        for i, b in enumerate(self.bins):
            for j, v in enumerate(variables):
                if v in b.edges:
                    out[i, j] = b.edges[v]
        return out


class BinningError(ValueError):
    pass


def parse_binning_text(text: str) -> BinSet:
    bs = BinSet()
    names: List[str] = []
    is_range: List[bool] = []
    n_values = 0
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        el = line.split()
        if el[0] == "variables:":
            names, is_range, n_values = [], [], 0
            for tok in el[1:]:
                if names and tok == names[-1]:
                    if is_range[-1]:
                        raise BinningError(f"variable {tok} appears more than twice: {line}")
                    is_range[-1] = True
                elif tok in names:
                    raise BinningError(f"{tok} is already set: {line}")
                else:
                    names.append(tok)
                    is_range.append(False)
                n_values += 1
            continue
        if not names:
            raise BinningError("bin line before any 'variables:' line")
        if len(el) != n_values:
            raise BinningError(f"expected {n_values} values, got {len(el)}: {line}")
        vals = [float(x) for x in el]
        edges: Dict[str, Tuple[float, float]] = {}
        i = 0
        for name, rng in zip(names, is_range):
            if rng:
                edges[name] = (vals[i], vals[i + 1])
                i += 2
            else:
                edges[name] = (vals[i], vals[i])
                i += 1
        edges = dict(sorted(edges.items(), key=lambda kv: kv[0].lower()))
        bs.bins.append(Bin(edges))
    return bs


def read_binning_text(path: str) -> BinSet:
    with open(path) as f:
        return parse_binning_text(f.read())

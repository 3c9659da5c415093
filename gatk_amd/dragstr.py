"""DRAGstr input side: STR context of read positions and per-position GOP/GCP imputation.

Mirrors the reference's classes for this step:
  * ``DragstrParams`` -- the (period x repeats) GOP / GCP / API tables, the lookup that clamps period
    and repeats to the table (DragstrParams.java:194-230), the built-in defaults (:49-128) and the
    text file format (DragstrParamUtils.java:32-60, :117-200);
  * ``DragstrReadSTRAnalyzer`` -- best period and repeat count per read position
    (DragstrReadSTRAnalyzer.java:69-139);
  * ``DragstrPairHMMInputScoreImputator`` -- the per-read GOP / GCP bytes the PairHMM receives
    (DragstrPairHMMInputScoreImputator.java:45-88).

The analysis and imputation run in the native library (``pairhmm_dragstr_*``); ``contexts`` gives
the table cell every read position reads, which is what the penalty fit differentiates through.
"""
from __future__ import annotations

import io
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import numpy as np

from . import _native
from .pairhmm import _raise_for

DEFAULT_MAX_PERIOD = 8
DEFAULT_MAX_REPEATS = 20

# DragstrParams.java:49-60 (GOP) and :79-90 (API); GCP = round(1000 / period) / 100 (:114-117).
_DEFAULT_GOP = [
    [45.00, 45.00, 45.00, 45.00, 45.00, 45.00, 40.50, 33.50, 28.00, 24.00, 21.75, 21.75, 21.75, 21.75, 21.75, 21.75, 21.75, 21.75, 21.75, 21.75],
    [39.50, 39.50, 39.50, 39.50, 36.00, 30.00, 27.25, 25.00, 24.25, 24.75, 26.25, 26.25, 26.25, 26.25, 26.25, 26.25, 26.25, 26.25, 26.25, 26.75],
    [38.50, 41.00, 41.00, 41.00, 41.00, 37.50, 35.25, 34.75, 34.75, 33.25, 33.25, 33.25, 32.50, 30.75, 28.50, 29.00, 29.00, 29.00, 29.00, 29.00],
    [37.50, 39.00, 39.00, 37.75, 34.00, 34.00, 30.25, 30.25, 30.25, 30.25, 30.25, 30.25, 30.25, 30.25, 30.25, 31.75, 31.75, 31.75, 31.75, 31.75],
    [37.00, 40.00, 40.00, 40.00, 36.00, 35.00, 24.50, 24.50, 24.50, 24.50, 22.50, 22.50, 22.50, 23.50, 23.50, 23.50, 23.50, 23.50, 23.50, 23.50],
    [36.25, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00, 40.00],
    [36.00, 40.50, 40.50, 40.50, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75, 20.75],
    [36.25, 39.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75, 32.75]]
_DEFAULT_API = [
    [39.00, 39.00, 37.00, 35.00, 32.00, 26.00, 20.00, 16.00, 12.00, 10.00, 8.00, 7.00, 7.00, 6.00, 6.00, 5.00, 5.00, 4.00, 4.00, 4.00],
    [30.00, 30.00, 29.00, 22.00, 17.00, 14.00, 11.00, 8.00, 6.00, 5.00, 4.00, 4.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 2.00, 2.00],
    [27.00, 27.00, 25.00, 18.00, 14.00, 12.00, 9.00, 7.00, 5.00, 4.00, 3.00, 3.00, 3.00, 3.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00],
    [27.00, 27.00, 18.00, 9.00, 9.00, 9.00, 9.00, 3.00, 3.00, 3.00, 3.00, 3.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00],
    [29.00, 29.00, 18.00, 8.00, 8.00, 8.00, 4.00, 3.00, 3.00, 3.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00, 2.00],
    [25.00, 25.00, 10.00, 10.00, 10.00, 4.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00, 3.00],
    [21.00, 21.00, 11.00, 11.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00, 5.00],
    [18.00, 18.00, 10.00, 6.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00, 4.00]]


class BadInput(ValueError):
    """UserException.BadInput for malformed parameter tables or files."""


@dataclass
class DragstrParams:
    gop_table: np.ndarray  # [max_period, max_repeats] Phred
    gcp_table: np.ndarray
    api_table: np.ndarray
    name: str = "<non-persistent>"

    def __post_init__(self) -> None:
        for nm in ("gop_table", "gcp_table", "api_table"):
            t = np.ascontiguousarray(np.asarray(getattr(self, nm), dtype=np.float64))
            if t.ndim != 2 or t.shape != np.asarray(self.gop_table).shape:
                raise BadInput(f"the {nm[:3].upper()} matrix has the wrong shape")
            if not np.all(np.isfinite(t) & (t >= 0)):
                raise BadInput(f"bad {nm[:3]} value: tables hold finite values >= 0")
            setattr(self, nm, t)

    @property
    def max_period(self) -> int:
        return int(self.gop_table.shape[0])

    @property
    def max_repeats(self) -> int:
        return int(self.gop_table.shape[1])

    def _lookup(self, table: np.ndarray, period: int, repeats: int) -> float:
        if period < 1 or repeats < 1:
            raise ValueError("period and repeats must be positive")
        return float(table[min(period, self.max_period) - 1, min(repeats, self.max_repeats) - 1])

    def gop(self, period: int, repeats: int) -> float:
        return self._lookup(self.gop_table, period, repeats)

    def gcp(self, period: int, repeats: int) -> float:
        return self._lookup(self.gcp_table, period, repeats)

    def api(self, period: int, repeats: int) -> float:
        return self._lookup(self.api_table, period, repeats)

    @classmethod
    def default(cls) -> "DragstrParams":
        gcp = np.array([[round(1000.0 / p) / 100.0] * DEFAULT_MAX_REPEATS for p in range(1, DEFAULT_MAX_PERIOD + 1)])
        return cls(np.array(_DEFAULT_GOP), gcp, np.array(_DEFAULT_API), name="<default>")

    # -- text format (DragstrParamUtils.java:32-60) -------------------------------------------------
    @classmethod
    def parse(cls, text: str, name: str = "<text>") -> "DragstrParams":
        lines = [ln for ln in text.splitlines()]
        k = 0
        while k < len(lines) and lines[k].startswith("#"):
            k += 1
        if k == len(lines):
            raise BadInput(f"there is no content in the dragstr-params file {name}")
        try:
            header = [int(x) for x in lines[k].split()]
        except ValueError as e:
            raise BadInput("bad format for an integer") from e
        if header != list(range(1, len(header) + 1)):
            raise BadInput(f"the DRAGstr parameter file header line must contain integers starting at 1 {header}")
        tables: Dict[str, list] = {}
        tname: Optional[str] = None
        rows: list = []
        for ln in lines[k + 1:]:
            if not ln.strip():
                continue
            if ln.rstrip().endswith(":"):
                if tname is not None:
                    tables[tname] = rows
                tname, rows = ln.rstrip()[:-1], []
                continue
            try:
                vals = [float(x) for x in ln.split()]
            except ValueError as e:
                raise BadInput(f"bad value in table {tname}") from e
            if len(vals) != len(header):
                raise BadInput(f"the {tname} matrix contains rows with length that does not match the max repeat length")
            rows.append(vals)
        if tname is None or tname == "":
            raise BadInput("table with no name")
        tables[tname] = rows
        for t in ("GOP", "GCP", "API"):
            if t not in tables:
                raise BadInput(f"missing table {t}")
        return cls(np.array(tables["GOP"]), np.array(tables["GCP"]), np.array(tables["API"]), name=name)

    @classmethod
    def read(cls, path: str) -> "DragstrParams":
        with open(path) as f:
            return cls.parse(f.read(), name=path)

    def format(self, **annotations) -> str:
        out = io.StringIO()
        out.write("#" * 92 + "\n# DragstrParams\n# -------------------------\n")
        for k, v in annotations.items():
            out.write(f"# {k}" + (f" = {v}" if v is not None else "") + "\n")
        out.write("#" * 92 + "\n")
        out.write("  ".join(f"{j:>5d}" for j in range(1, self.max_repeats + 1)) + "\n")
        for nm, t in (("GOP", self.gop_table), ("GCP", self.gcp_table), ("API", self.api_table)):
            out.write(nm + ":\n")
            for row in t:
                out.write("  ".join(f"{v:5.2f}" for v in row) + "\n")
        return out.getvalue()

    def write(self, path: str, **annotations) -> None:
        with open(path, "w") as f:
            f.write(self.format(**annotations))


class DragstrReadSTRAnalyzer:
    """Best period and repeat count per position of one read (DragstrReadSTRAnalyzer.java:69-139)."""

    def __init__(self, bases, max_period: int) -> None:
        if max_period < 1:
            raise ValueError("the input max period must be 1 or greater")
        b = np.ascontiguousarray(np.frombuffer(bytes(bases), dtype=np.uint8) if isinstance(bases, (bytes, bytearray))
                                 else np.asarray(bases, dtype=np.uint8))
        self.length = len(b)
        self.max_period = max_period
        self.period = np.zeros(self.length, dtype=np.int32)
        self.repeats = np.zeros(self.length, dtype=np.int32)
        lib = _native.load_library()
        _raise_for(lib.pairhmm_dragstr_analyze(b.ctypes.data, self.length, max_period, self.period.ctypes.data,
                                               self.repeats.ctypes.data), "dragstr_analyze")

    def _check(self, position: int) -> None:
        if not 0 <= position < self.length:
            raise IndexError(f"cannot query outside requested boundaries [0, {self.length}): {position}")

    def most_repeated_period(self, position: int) -> int:
        self._check(position)
        return int(self.period[position])

    def number_of_most_repeats(self, position: int) -> int:
        self._check(position)
        return int(self.repeats[position])


def _read_arrays(batch):
    rb = np.ascontiguousarray(batch.read_bases, dtype=np.uint8)
    ro = np.ascontiguousarray(batch.read_off, dtype=np.int64)
    rl = np.ascontiguousarray(batch.read_len, dtype=np.int32)
    return rb, ro, rl


def contexts(batch, max_period: int, max_repeats: int) -> np.ndarray:
    """Table cell (period-1)*max_repeats + repeats-1 of every read position of a packed batch;
    -1 at the last position of each read (fixed 45 / 10 there)."""
    rb, ro, rl = _read_arrays(batch)
    out = np.full(len(rb), -1, dtype=np.int32)
    lib = _native.load_library()
    _raise_for(lib.pairhmm_dragstr_contexts(rb.ctypes.data, ro.ctypes.data, rl.ctypes.data, len(rl), len(rb),
                                            max_period, max_repeats, out.ctypes.data), "dragstr_contexts")
    return out


class DragstrPairHMMInputScoreImputator:
    """GOP (insertion = deletion) and GCP bytes per read position from DRAGstr tables
    (DragstrPairHMMInputScoreImputator.java:53-70)."""

    def __init__(self, params: DragstrParams) -> None:
        self.params = params

    def impute_batch(self, batch) -> Tuple[np.ndarray, np.ndarray]:
        rb, ro, rl = _read_arrays(batch)
        gop = np.zeros(len(rb), dtype=np.uint8)
        gcp = np.zeros(len(rb), dtype=np.uint8)
        p = self.params
        lib = _native.load_library()
        _raise_for(lib.pairhmm_dragstr_impute(p.gop_table.ctypes.data, p.gcp_table.ctypes.data, p.max_period,
                                              p.max_repeats, rb.ctypes.data, ro.ctypes.data, rl.ctypes.data, len(rl),
                                              len(rb), gop.ctypes.data, gcp.ctypes.data), "dragstr_impute")
        return gop, gcp

    def impute(self, bases) -> Tuple[bytes, bytes]:
        """One read: (gop bytes used for insertions and deletions, gcp bytes)."""
        b = np.frombuffer(bytes(bases), dtype=np.uint8)

        class _One:
            read_bases = b
            read_off = np.zeros(1, dtype=np.int64)
            read_len = np.array([len(b)], dtype=np.int32)

        gop, gcp = self.impute_batch(_One)
        return gop.tobytes(), gcp.tobytes()

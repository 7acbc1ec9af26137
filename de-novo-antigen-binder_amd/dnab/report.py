"""Candidate reports (src/report.rs): the summary of one designed candidate
and its FASTA, JSON and CSV writers, formatted by the native library
(dnab_write_report) as the reference formats them. Text out; the caller
writes the file (src/main.rs:516-530)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

from . import _native as N

FASTA, JSON, CSV = 0, 1, 2


@dataclass
class CandidateReport:
    """CandidateReport (src/report.rs:8-20); light_sequence None for a
    single chain, clashes / rama_outliers None when not computed."""
    rank: int
    sequence: str
    light_sequence: Optional[str]
    energy: float
    entropy_penalty: float
    dg_corrected: float
    net_charge: int
    aggregation_risk: bool
    n_interface: int
    clashes: Optional[int] = None
    rama_outliers: Optional[int] = None


def _write(reports: list[CandidateReport], fmt: int) -> str:
    arr = (N.CandidateReportC * max(len(reports), 1))()
    keep = []
    for k, r in enumerate(reports):
        seq = r.sequence.encode()
        light = r.light_sequence.encode() if r.light_sequence is not None else None
        if b"\0" in seq or (light and b"\0" in light):
            raise ValueError("report sequences cannot hold NUL")
        keep.append((seq, light))
        arr[k] = N.CandidateReportC(
            r.rank, seq, light, r.energy, r.entropy_penalty, r.dg_corrected, r.net_charge,
            1 if r.aggregation_risk else 0, r.n_interface,
            -1 if r.clashes is None else r.clashes,
            -1 if r.rama_outliers is None else r.rama_outliers)
    n = C.c_size_t(0)
    N.check(N.lib.dnab_write_report(arr, len(reports), fmt, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    N.check(N.lib.dnab_write_report(arr, len(reports), fmt, buf, n.value + 1, C.byref(n)))
    return buf.raw[:n.value].decode()


def write_fasta(reports: list[CandidateReport]) -> str:
    """write_fasta (src/report.rs:41-56): one record per chain."""
    return _write(reports, FASTA)


def write_json(reports: list[CandidateReport]) -> str:
    """write_json (src/report.rs:58-89): a JSON array, floats to 6 places."""
    return _write(reports, JSON)


def write_csv(reports: list[CandidateReport]) -> str:
    """write_csv (src/report.rs:91-113): header plus one row per candidate."""
    return _write(reports, CSV)

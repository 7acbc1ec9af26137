"""Candidate report writers (src/report.rs:41-113) through dnab_write_report.

Pinned by the reference's own tests (src/report.rs:135-171) and by reading
the text back: json.loads and the csv module recover every field (escaped
quotes, backslashes and newlines included), FASTA splits into its records;
floats are the f32 value rounded half to even at the stated places (exact
binary ties chosen so that rounding half up would differ), and non-finite
values print as Rust prints them (NaN, inf, -inf)."""
import csv
import io
import json

import numpy as np
import pytest

from dnab.report import CandidateReport, write_csv, write_fasta, write_json


def sample(**kw):
    r = CandidateReport(rank=1, sequence="ACDEFG", light_sequence=None, energy=-12.5,
                        entropy_penalty=5.7, dg_corrected=-6.8, net_charge=2,
                        aggregation_risk=False, n_interface=4, clashes=0, rama_outliers=1)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


# the reference's tests
def test_fasta_contains_header_and_sequence():
    out = write_fasta([sample()])
    assert ">candidate_1" in out and "ACDEFG" in out


def test_fasta_emits_two_records_for_fv_candidates():
    out = write_fasta([sample(light_sequence="HIKLMN")])
    assert "_VH" in out and "_VL" in out and "HIKLMN" in out


def test_json_is_well_formed_brackets():
    out = write_json([sample(), sample()])
    assert out.startswith("[") and out.rstrip().endswith("]")
    assert out.count("{") == 2 and out.count("}") == 2


def test_csv_has_header_plus_one_row_per_candidate():
    lines = write_csv([sample(), sample()]).splitlines()
    assert len(lines) == 3 and lines[0].startswith("rank,")


# read back
def test_fasta_records():
    rs = [sample(rank=3, energy=-7.25), sample(rank=4, light_sequence="QQ", energy=1.0)]
    assert write_fasta(rs) == (">candidate_3 E=-7.2500\nACDEFG\n"
                               ">candidate_4_VH E=1.0000\nACDEFG\n>candidate_4_VL E=1.0000\nQQ\n")


def test_json_reads_back():
    odd = 'A"B\\C\nD'
    rs = [sample(rank=2, sequence=odd, light_sequence="EV\"IL", net_charge=-3,
                 aggregation_risk=True, clashes=None, rama_outliers=None),
          sample(rank=5, n_interface=17)]
    got = json.loads(write_json(rs))
    assert [g["rank"] for g in got] == [2, 5]
    assert got[0]["sequence"] == odd and got[0]["light_sequence"] == 'EV"IL'
    assert got[1]["light_sequence"] is None
    assert got[0]["net_charge"] == -3 and got[0]["aggregation_risk"] is True
    assert got[0]["clashes"] is None and got[0]["rama_outliers"] is None
    assert got[1]["clashes"] == 0 and got[1]["rama_outliers"] == 1 and got[1]["n_interface"] == 17
    for g, r in zip(got, rs):
        for k in ("energy", "entropy_penalty", "dg_corrected"):
            assert g[k] == round(float(np.float32(getattr(r, k))), 6)
    assert list(got[0]) == ["rank", "sequence", "light_sequence", "energy", "entropy_penalty",
                            "dg_corrected", "net_charge", "aggregation_risk", "n_interface",
                            "clashes", "rama_outliers"]


def test_csv_reads_back():
    rs = [sample(sequence='Q,"W"', light_sequence=None, clashes=None),
          sample(rank=9, light_sequence="L,L", rama_outliers=None, aggregation_risk=True)]
    rows = list(csv.reader(io.StringIO(write_csv(rs))))
    assert rows[0] == ["rank", "sequence", "light_sequence", "energy", "entropy_penalty",
                       "dg_corrected", "net_charge", "aggregation_risk", "n_interface", "clashes",
                       "rama_outliers"]
    assert rows[1] == ["1", 'Q,"W"', "", "-12.500000", "5.700000", "-6.800000", "2", "false", "4",
                       "", "1"]
    assert rows[2][0] == "9" and rows[2][2] == "L,L" and rows[2][7] == "true"
    assert rows[2][9] == "0" and rows[2][10] == ""


def test_rounding_and_special_values():
    # 1/32 and 3/32 are exact f32 ties at 4 places: half to even gives ...12 and ...38
    assert write_fasta([sample(energy=0.03125)]).startswith(">candidate_1 E=0.0312\n")
    assert write_fasta([sample(energy=-0.09375)]).startswith(">candidate_1 E=-0.0938\n")
    assert "E=NaN" in write_fasta([sample(energy=float("nan"))])
    assert '"energy": inf,' in write_json([sample(energy=float("inf"))])
    row = write_csv([sample(entropy_penalty=float("-inf"), dg_corrected=-0.0)]).splitlines()[1]
    assert ",-inf,-0.000000," in row
    # the f32 value, not the double the caller passed
    assert "E=0.1000" in write_fasta([sample(energy=0.1)])
    assert write_csv([sample(energy=1e-7)]).splitlines()[1].split(",")[3] == "0.000000"


def test_empty_and_bad_input():
    assert write_fasta([]) == ""
    assert write_json([]) == "[\n]\n"
    assert write_csv([]).count("\n") == 1
    with pytest.raises(ValueError):
        write_fasta([sample(sequence="A\0B")])

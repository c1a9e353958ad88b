"""The golden text format (tests/golden_io.hpp:12-81 of the reference): the package writer
reproduces the reference-written fixtures byte for byte, the reader agrees with the checkers'
reader, and malformed files are rejected like the reference reader does."""
import pathlib

import numpy as np
import pytest

from checkers import read_golden as checker_read
from mhd_gpu_solver_amd import state_io

GOLD = pathlib.Path(__file__).parent / "golden"
FIXTURES = sorted(GOLD.glob("*.golden"))


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.name)
def test_round_trip_is_byte_identical(path, tmp_path):
    snap = state_io.read_golden(path)
    out = tmp_path / path.name
    state_io.write_golden(out, snap)
    assert out.read_bytes() == path.read_bytes()


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.name)
def test_reader_agrees_with_checkers(path):
    snap = state_io.read_golden(path)
    meta, fields = checker_read(path)
    assert snap.meta == meta
    assert snap.fields.keys() == fields.keys()
    for k in fields:
        assert np.array_equal(snap.fields[k], fields[k])


def test_sorted_output_and_values(tmp_path):
    snap = state_io.GoldenSnapshot({"z": "1", "a": "x"},
                                   {"by": np.array([-0.0, 1.5]), "ax": np.array([1e-300])})
    state_io.write_golden(tmp_path / "s.golden", snap)
    text = (tmp_path / "s.golden").read_text().splitlines()
    assert text == ["MHD_GOLDEN 1", "a x", "z 1", "field ax 1", "1.00000000000000003e-300",
                    "field by 2", "-0.00000000000000000e+00", "1.50000000000000000e+00"]
    back = state_io.read_golden(tmp_path / "s.golden")
    assert back.meta == snap.meta and np.array_equal(back.fields["by"], snap.fields["by"])


@pytest.mark.parametrize("text", ["MHD_GOLD 1\n", "MHD_GOLDEN 2\n", "",
                                  "MHD_GOLDEN 1\nfield rho 3\n1.0\n2.0\n",
                                  "MHD_GOLDEN 1\nsteps\n"])
def test_malformed_files_rejected(tmp_path, text):
    p = tmp_path / "bad.golden"
    p.write_text(text)
    with pytest.raises(ValueError):
        state_io.read_golden(p)


def test_cpp_to_string():
    assert state_io.cpp_to_string(36) == "36"
    assert state_io.cpp_to_string(1.7136041999) == "1.713604"
    assert state_io.cpp_to_string(0.1) == "0.100000"

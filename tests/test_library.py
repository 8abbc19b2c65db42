"""CPU tests of the product library without a GPU: it loads, exports every symbol the C header
declares, and its host-only schedule agrees with the oracle's sweep count."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import lib as H


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(H.LIB_PATH)
    names = H.exported_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(H.exported_symbols()) == set(H.SIGNATURES)


@pytest.mark.parametrize("levels,min_level,D", [(9, 2, 64), (11, 2, 256), (6, 1, 16), (12, 7, 256), (3, 1, 5)])
def test_schedule_matches_oracle(levels, min_level, D):
    lib = H.load()
    assert lib.hbp_schedule_sweeps(levels, min_level, D) == O.lib.orc_bp_total_sweeps(levels, min_level, D)


def test_schedule_counts_by_hand():
    # L = 3, min_level = 1: one BP level (l = 1), iter_bp = 20 outer iterations of 5 sweeps,
    # plus a random sweep for i > 15 while max_dz = D/2 * 1/3 >= 0.01 halves: 4 of them
    lib = H.load()
    assert lib.hbp_schedule_sweeps(3, 1, 64) == 20 * 5 + 4


def test_errors_are_reported_not_raised_into_the_caller():
    lib = H.load()
    assert lib.hbp_bp_step(None, None, 1, None) != 0
    assert b"NULL" in lib.hbp_last_error()

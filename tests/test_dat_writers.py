"""The reference's .dat writers (modGlobalMetod.f90:186-306) as the C ABI
writes them, checked byte for byte against
  - what the flang build of the reference's own writers wrote for values
    that probe the F16.8 and I10 fields (signed zeros, negative values that
    round to zero, decimal ties, the field-width limits, NaN, infinities);
  - .dat files the reference's own runs wrote (Release_Intel64): a current
    from guardarVectorC and two monostatic tables from guardarMatrizR, read
    back and written again.
Host code only."""
import os

import numpy as np
import pytest

import lywell_amd as lw

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dat_writers.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def read(path):
    with open(path) as f:
        return f.read()


def test_vector_c_matches_the_reference_writer(gold, tmp_path):
    v = gold["vals"]
    c = np.empty(v.size // 2, np.complex128)  # built part by part: keeps -0.0 and NaN as they are
    c.real = v[0::2][: v.size // 2]
    c.imag = v[1::2][: v.size // 2]
    lw.write_dat_vector_c(tmp_path / "vc.dat", c)
    assert read(tmp_path / "vc.dat") == str(gold["flang_vc"])


def test_matrix_r_matches_the_reference_writer(gold, tmp_path):
    v = gold["vals"]
    m = v[: 3 * (v.size // 3)].reshape(3, -1).T  # reshape(v, [n/3, 3]) in Fortran order
    lw.write_dat_matrix_r(tmp_path / "mr.dat", m)
    assert read(tmp_path / "mr.dat") == str(gold["flang_mr"])


def test_vector_r_and_i_match_the_reference_writers(gold, tmp_path):
    lw.write_dat_vector_r(tmp_path / "vr.dat", gold["vals"])
    lw.write_dat_vector_i(tmp_path / "vi.dat", gold["iv"])
    assert read(tmp_path / "vr.dat") == str(gold["flang_vr"])
    assert read(tmp_path / "vi.dat") == str(gold["flang_vi"])


@pytest.mark.parametrize("key,kind", [("ifort_current", "c"), ("ifort_mono_vv", "r"),
                                      ("ifort_mono_incremental", "r")])
def test_round_trip_of_files_the_reference_wrote(gold, tmp_path, key, kind):
    txt = str(gold[key])
    a = np.loadtxt(txt.splitlines(), ndmin=2)
    out = tmp_path / "o.dat"
    if kind == "c":
        c = np.empty(len(a), np.complex128)
        c.real, c.imag = a[:, 0], a[:, 1]
        lw.write_dat_vector_c(out, c)
    else:
        lw.write_dat_matrix_r(out, a)
    assert read(out) == txt


def test_field_width_limits():
    def f(v):
        b = np.zeros(17, np.uint8)
        assert lw.lib().lw_dat_format_f16_8(lw._C.c_double(v), lw._P(b)) == 0
        return bytes(b[:16]).decode()
    assert f(9999999.99999999) == "9999999.99999999"
    assert f(-999999.99999999) == "-999999.99999999"
    assert f(-9999999.0) == "*" * 16
    assert f(-1e-10) == "     -0.00000000"

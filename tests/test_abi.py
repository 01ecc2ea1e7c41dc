"""The C-ABI library: loads, exports every function include/fsm.h declares,
and its host-side parameter logic agrees with the reference restatement
(no GPU needed)."""
import ctypes
import os

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = fsm_amd.lib()
    names = fsm_amd.exported_symbols_from_header(os.path.join(ROOT, "include", "fsm.h"))
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert fsm_amd.version().startswith("fsm-mi355x")


def test_params_struct_matches_reference_layout():
    # struct parameters, src/Parameter.h:2-33: 28 four-byte fields
    assert ctypes.sizeof(fsm_amd.Params) == 28 * 4
    assert [f for f, _ in fsm_amd.Params._fields_][:3] == ["OUTPUTFORMAT", "BARYON", "IC_ENERGY"]
    assert fsm_amd.Params._fields_[-2][0] == "DIM" and fsm_amd.Params._fields_[-1][0] == "TAU"


def test_defaults_match_reference():
    a = fsm_amd.params_default().as_dict()
    b = O.params_default().as_dict()
    for k in ("E_R", "TAU_R", "ALPHA"):  # no reference default
        a.pop(k), b.pop(k)
    assert a == b
    assert a["DIM_X"] == 101 and a["DIM_PHIP"] == 51 and a["IC_ENERGY"] == 5


def test_positional_reader_on_reference_input_file():
    path = os.path.join(ROOT, "tests", "golden", "freestream_input.reference")
    p, st = fsm_amd.params_read(path)
    assert st == 0
    assert (p.OUTPUTFORMAT, p.BARYON, p.IC_ENERGY, p.IC_BARYON) == (1, 0, 1, 1)
    assert (p.DIM_X, p.DIM_Y, p.DIM_ETA, p.DIM_RAP, p.DIM_PHIP) == (201, 201, 1, 1, 500)
    assert p.DX == pytest.approx(0.1) and p.DRAP == pytest.approx(0.07) and p.DTAU == pytest.approx(1.0)
    assert p.TAU0 == 0.0 and p.E_FREEZE == pytest.approx(1.7) and p.E_DEP_FS == 1
    assert p.E_R == pytest.approx(4000.0) and p.TAU_R == pytest.approx(1.2) and p.ALPHA == pytest.approx(2.5)


def test_missing_parameter_file_keeps_defaults():
    p, st = fsm_amd.params_read("/nonexistent/freestream_input")
    assert st == fsm_amd.FSM_ERR_FILE
    assert p.as_dict() == fsm_amd.params_default().as_dict()


@pytest.mark.parametrize("name", ["C0", "C1", "C2"])
def test_finalize_matches_reference_overrides(name):
    p, e = W.config(name)
    a = fsm_amd.params_finalize(p, e)
    b = O.params_finalize(p, e)
    assert a.as_dict() == b.as_dict()
    if name == "C2":
        assert a.DIM_PHIP == 51  # 48 raised by ceil(2 pi DTAU / min(DX,DY))


def test_energy_dependent_tau_matches_reference():
    p, e = W.config("C0")
    p.E_DEP_FS, p.E_R, p.TAU_R, p.ALPHA = 1, 4.0, 1.2, 0.25
    a = fsm_amd.params_finalize(p, e)
    b = O.params_finalize(p, e)
    assert a.DTAU == b.DTAU and a.TAU == b.TAU and a.DIM_PHIP == b.DIM_PHIP
    assert a.DTAU != p.DTAU


def test_size_mismatch_is_reported():
    p, e = W.config("C0")
    with pytest.raises(fsm_amd.FsmError) as ei:
        fsm_amd.params_finalize(p, e[:-1])
    assert ei.value.status == fsm_amd.FSM_ERR_SIZE

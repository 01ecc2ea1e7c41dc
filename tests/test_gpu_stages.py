"""The function-level boundary: the reference's stage functions under their
own names and signatures (freestream-milne_amd/csrc/fsm_stages.h over the
fsm_solve_eigen_system / fsm_calculate_* entry points), called on float**
arrays in the reference's order (src/FreestreamMilne.cpp:443-506) by the
fsm_stages program, on the oracle's own float stress tensor.

  * solveEigenSystem (src/LandauMatch.cpp:145-331): e, u against the
    oracle's FP64 eigen-solve of the same tensor with the reference's
    selection rules, on every cell with e > 1e-3 e_max, and the same
    fallback decision on every cell;
  * calculatePressure, calculateBulkPressure, calculateShearViscTensor
    (src/EquationOfState.cpp:5-80, src/LandauMatch.cpp:332-367): bitwise the
    reference float path's arithmetic on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O

pytestmark = pytest.mark.gpu

FSM_STAGES = os.path.join(os.path.dirname(fsm_amd.LIB_PATH), "fsm_stages")


def _case(name):
    if name == "small3d":
        p, e = W.small_3d()
    else:
        p, e = W.config("C1" if name.startswith("C1") else "C0")
    if name.endswith("eos2"):
        p.EOS_TYPE = 2
    return fsm_amd.params_finalize(p, e), e


@pytest.mark.parametrize("name", ["C1", "C0_eos2", "small3d"])
def test_reference_stage_functions(name, tmp_path):
    assert os.path.exists(FSM_STAGES), "fsm_stages is built by __graft_entry__.build()"
    q, e = _case(name)
    ref = O.run(q, e, "float")
    T = ref["T"].astype(np.float32)
    D = q.DIM
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(bytes(q))
        f.write(np.ascontiguousarray(T).tobytes())
    print("fsm_stages", name, "DIM", D, flush=True)
    r = subprocess.run([FSM_STAGES, "in.bin", "out.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float32).reshape(17, D)
    eg, ug, pg, Pig, pig = out[0], out[1:5], out[5], out[6], out[7:17]
    # solveEigenSystem against the oracle's FP64 solve of the same tensor
    eo, uo = O.landau(T.astype(np.float64).T, np.float32(q.TAU))
    m = eo > 1e-3 * eo.max()
    ee = (np.abs(eg - eo)[m] / eo[m]).max()
    ue = (np.abs(ug.T - uo)[m].max(axis=1) / uo[m, 0]).max()
    fo = (eo == np.float64(np.float32(1e-3))) & (uo[:, 0] == 1.0)
    fg = (eg == np.float32(1e-3)) & (ug[0] == 1.0)
    print(name, "solveEigenSystem: e %.3g, u %.3g, fallbacks %d, mismatches %d" %
          (ee, ue, int(fo.sum()), int((fo != fg).sum())), flush=True)
    assert ee < 1e-5 and ue < 1e-5
    assert np.array_equal(fo, fg)
    # the per-cell stages: the reference float arithmetic on the same inputs
    assert np.array_equal(pg, O.pressure(eg, q.EOS_TYPE, "float").astype(np.float32))
    Pio, pio = O.viscous_float(q, T, eg, ug, pg)
    assert np.array_equal(Pig, Pio)
    assert np.array_equal(pig, pio)

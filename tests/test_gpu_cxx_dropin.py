"""The C++ drop-in, run as a C++ caller runs it: class FREESTREAMMILNE of
freestream-milne_amd/csrc/FreestreamMilne.h, driven by the fsm_run program
(csrc/run_wrapper.cpp) the way src/RunWrapper.cpp:3-43 drives the reference's
class -- ./freestream_input in the working directory (read positionally,
src/FileIO.cpp:166-245), initialize_from_vector, run_freestream_milne(),
output_to_vectors (src/FreestreamMilne.cpp:97-136).  All 17 vectors against
the oracle's reference float path, with the tolerances of tests/parity.py, on
C0 and C1; and against the Python mirror of the same class."""
import os
import subprocess

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W
import oracle as O
from parity import TOL, output_errs, well_conditioned

pytestmark = pytest.mark.gpu

FSM_RUN = os.path.join(os.path.dirname(fsm_amd.LIB_PATH), "fsm_run")

# the order src/FileIO.cpp:185-240 reads the fields in (names are ignored)
ORDER = ["OUTPUTFORMAT", "BARYON", "IC_ENERGY", "IC_BARYON", "ETA_WIDTH", "ETA_FLAT", "SIGMA", "SIGMA_B", "DIM_X",
         "DIM_Y", "DIM_ETA", "DIM_RAP", "DIM_PHIP", "DX", "DY", "DETA", "DRAP", "DTAU", "TAU0", "EOS_TYPE",
         "E_FREEZE", "VISCOUS_MATCHING", "E_DEP_FS", "E_R", "TAU_R", "ALPHA"]


def write_freestream_input(path, p):
    lines = []
    for k in ORDER:
        v = getattr(p, k)
        lines.append("%s\t%s" % (k, ("%d" % v) if isinstance(v, int) else ("%.9g" % v)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("name", ["C0", "C1"])
def test_cxx_class_through_fsm_run(name, tmp_path):
    assert os.path.exists(FSM_RUN), "fsm_run is built by __graft_entry__.build()"
    p, e = W.config(name)
    write_freestream_input(tmp_path / "freestream_input", p)
    np.savetxt(tmp_path / "profile.txt", e.astype(np.float32), fmt="%.9g")
    print("fsm_run", name, flush=True)
    r = subprocess.run([FSM_RUN, "profile.txt", "out.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-400:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "run_freestream_milne() ran sucessfully" in r.stdout
    q = fsm_amd.params_finalize(p, e)
    tau = float(r.stdout.split("tau_LandauMatch = ")[1].split()[0])
    assert abs(tau - q.TAU) < 1e-6
    got = np.loadtxt(tmp_path / "out.txt").T  # (17, DIM), %.9g
    assert got.shape == (17, q.DIM)
    ref = O.run(q, e, "float")
    errs = output_errs(got, ref["out"], np.abs(ref["T"][0]).max(), well_conditioned(ref["T"], q.TAU))
    print(name, "C++ class vs oracle:", errs, flush=True)
    for k in ("e", "p", "u", "pi", "Pi"):
        assert errs[k] <= TOL, (k, errs[k])
    # the same class in Python (same library, same kernels): equal to the
    # nine significant digits fsm_run prints
    fs = fsm_amd.FreestreamMilne(input_file=str(tmp_path / "freestream_input"), output_dir=None)
    fs.initialize_from_vector(e)
    assert fs.run_freestream_milne() == 0
    vecs = np.stack(fs.output_to_vectors())
    scale = np.maximum(np.abs(vecs), 1e-300)
    assert np.all((np.abs(got - vecs) <= 6e-9 * scale) | (np.abs(got - vecs) <= 1e-30))


def test_cxx_class_size_mismatch_exits_like_the_reference(tmp_path):
    """A vector whose size differs from DIM: the reference prints both sizes
    and calls exit(-1) (src/FreestreamMilne.cpp:259-264)."""
    p, e = W.config("C0")
    write_freestream_input(tmp_path / "freestream_input", p)
    np.savetxt(tmp_path / "profile.txt", e[:-7], fmt="%.9g")
    r = subprocess.run([FSM_RUN, "profile.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert "does not match freestream_input" in r.stdout and "Vector size : %d" % (e.size - 7) in r.stdout

"""HDF5 event logs (frictionqpotspringblock_amd/h5events.py) in the schema
the reference's examples read with h5py (examples/Line1d_Cuspy_Laplace.py:
63-67): a round trip through the library, and the file read by h5py itself
where an interpreter with h5py exists (the image's conda Python)."""

import os
import subprocess

import numpy as np
import pytest

from tests import protocols as P

H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _records(oracle, n=21):
    s = P.example_system(oracle, "Line1d_Cuspy_Laplace")
    u_frame, f_frame, S, inc = P.quasistatic(s, n)
    return {"u_frame": u_frame, "f_frame": f_frame, "S": S, "inc": inc}


def test_round_trip(oracle, tmp_path):
    from frictionqpotspringblock_amd import h5events
    from frictionqpotspringblock_amd.batch import load_events, save_events

    try:
        h5events.library()
    except RuntimeError as e:
        pytest.skip(str(e))
    q = _records(oracle)
    path = str(tmp_path / "events.h5")
    save_events(path, q)
    back = load_events(path)
    assert sorted(back) == ["S", "f_frame", "inc", "x_frame"]
    assert back["S"].dtype == np.int64 and back["x_frame"].dtype == np.float64
    assert np.array_equal(back["x_frame"], q["u_frame"])
    assert np.array_equal(back["f_frame"], q["f_frame"])
    assert np.array_equal(back["S"], q["S"])
    assert np.array_equal(back["inc"], q["inc"])
    # the golden prefix of the reference's own example file
    ref = P.golden("Line1d_Cuspy_Laplace")
    assert np.array_equal(back["S"], ref["S"][: len(back["S"])])
    # a batch's records: [R, nevents], one row per system
    qb = {k: np.stack([v, v]) for k, v in q.items()}
    save_events(path, qb)
    assert load_events(path)["S"].shape == (2, len(q["S"]))
    save_events(path, qb, system=1)
    assert np.array_equal(load_events(path)["S"], q["S"])


@pytest.mark.skipif(not os.path.exists(H5PY_PYTHON), reason="no interpreter with h5py")
def test_read_by_h5py(oracle, tmp_path):
    """The examples' own check (examples/Line1d_Cuspy_Laplace.py:64-67) run by
    h5py on a file written here."""
    from frictionqpotspringblock_amd import h5events
    from frictionqpotspringblock_amd.batch import save_events

    try:
        h5events.library()
    except RuntimeError as e:
        pytest.skip(str(e))
    q = _records(oracle)
    path = str(tmp_path / "events.h5")
    save_events(path, q)
    ref = tmp_path / "ref.npz"
    np.savez(ref, x_frame=q["u_frame"], f_frame=q["f_frame"], S=q["S"])
    code = (
        "import h5py, numpy as np\n"
        f"r = np.load({str(ref)!r})\n"
        f"with h5py.File({path!r}, 'r') as file:\n"
        "    assert np.allclose(r['x_frame'], file['x_frame'][...])\n"
        "    assert np.allclose(r['f_frame'], file['f_frame'][...])\n"
        "    assert np.all(r['S'] == file['S'][...])\n"
        "    assert file['S'].dtype == np.int64\n"
        "print('h5py ok')\n"
    )
    # a clean environment: this interpreter's variables (paths of another
    # Python and its libraries) must not leak into the other one
    env = {"PATH": "/usr/bin:/bin", "HOME": os.environ.get("HOME", "/tmp")}
    r = subprocess.run([H5PY_PYTHON, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py absent")
    assert r.returncode == 0, r.stderr
    assert "h5py ok" in r.stdout

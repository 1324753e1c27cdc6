"""CPU checks of the drop-in boundary: the C ABI library loads, exports every
function declared in include/fqsb.h, and the host layer raises the
reference's errors before touching a device."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fqsb.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fqsb_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from frictionqpotspringblock_amd import batch

    lib = ctypes.CDLL(batch.LIBRARY_PATH)
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python layer declares a signature for every exported entry point
    assert set(names) == set(batch.SIGNATURES)
    assert lib.fqsb_version  # callable without a device
    batch.capi()


def test_version_without_device():
    from frictionqpotspringblock_amd import batch

    assert b"gfx950" in batch.capi().fqsb_version()


def test_unknown_distribution_is_a_host_error():
    import frictionqpotspringblock_amd as fq

    with pytest.raises(RuntimeError, match="Unknown distribution: foo"):
        fq.Line1d.System_Cuspy_Laplace(
            m=1, eta=1, mu=1, k_interactions=1, k_frame=0.1, dt=1, shape=[5], seed=0,
            distribution="foo", parameters=[1.0],
        )


def test_api_surface_matches_reference_binding():
    """Class names and argument names of python/main.cpp:475-1006."""
    import frictionqpotspringblock_amd as fq

    line1d = [
        "System_Cuspy_Laplace",
        "System_Cuspy_Laplace_Nopassing",
        "System_Cuspy_Laplace_RandomForcing",
        "System_Cuspy_Quartic_RandomForcing",
        "System_SemiSmooth_Laplace",
        "System_Smooth_Laplace",
        "System_Cuspy_Quartic",
        "System_Cuspy_QuarticGradient",
        "System_Cuspy_LongRange",
    ]
    for name in line1d:
        assert hasattr(fq.Line1d, name)
    for name in ["System_Cuspy_Laplace", "System_Cuspy_QuarticGradient"]:
        assert hasattr(fq.Line2d, name)
    assert hasattr(fq.detail, "RandomNormalForcing_1")
    doc = fq.Line1d.System_Cuspy_LongRange.__init__.__doc__
    for arg in ["m", "eta", "mu", "k_interactions", "alpha", "k_frame", "dt", "shape", "seed", "distribution",
                "parameters", "offset", "nchunk"]:
        assert f"{arg}:" in doc
    assert "offset: typing.SupportsFloat | typing.SupportsIndex = -100.0" in doc
    S = fq.Line1d.System_Cuspy_Laplace
    for method in ["timeStep", "timeSteps", "timeStepsUntilEvent", "flowSteps", "minimise", "minimise_truncate",
                   "eventDrivenStep", "maxUniformDisplacement", "trigger", "advanceToFixedForce", "refresh",
                   "quench"]:
        assert hasattr(S, method), method
    for prop in ["u", "v", "a", "f", "f_potential", "f_frame", "f_interactions", "f_damping", "inc", "t", "u_frame",
                 "residual", "temperature", "chunk", "size", "shape", "dt", "mu", "eta", "m", "k_frame",
                 "quasistaticActivityFirst", "quasistaticActivityLast"]:
        assert hasattr(S, prop), prop
    assert not hasattr(fq.Line1d.System_Cuspy_Laplace_RandomForcing, "minimise")
    assert not hasattr(fq.Line1d.System_Cuspy_Laplace_Nopassing, "timeStep")
    assert hasattr(fq.Line1d.System_Cuspy_Laplace_RandomForcing, "external")


def _checkpoint(R, N, nchunk, external=False):
    import numpy as np

    n = R * N
    ck = {k: np.zeros(n) for k in ("u", "v", "a")}
    ck.update(
        y=np.tile(np.arange(nchunk, dtype=float), n),
        start=np.zeros(n, dtype=np.int64),
        jloc=np.full(n, nchunk // 2, dtype=np.int32),
        st_begin=np.zeros(n, dtype=np.uint64),
        st_end=np.zeros(n, dtype=np.uint64),
        inc=np.zeros(R, dtype=np.int64),
        u_frame=np.zeros(R),
        qs_first=np.zeros(R, dtype=np.int64),
        qs_last=np.zeros(R, dtype=np.int64),
    )
    if external:
        ck.update(f_thermal=np.zeros(n), next=np.zeros(n, dtype=np.int64), rng_state=np.zeros(R, dtype=np.uint64))
    return ck


@pytest.mark.parametrize("bad", [-1, 99, 4000])
def test_checkpoint_with_bad_well_index_is_refused_on_the_host(bad):
    """A checkpoint whose local well index lies outside the window (the device
    would read y[p][jloc + 1] out of the row) raises before any device call."""
    from frictionqpotspringblock_amd.batch import validate_checkpoint

    R, N, nchunk = 2, 5, 100
    ck = _checkpoint(R, N, nchunk)
    ck["jloc"][3] = nchunk - 2  # the last well of the window is valid
    validate_checkpoint(ck, R, [N], nchunk, False)
    ck["jloc"][7] = bad
    with pytest.raises(RuntimeError, match="local well index .* of particle 7 is outside the window"):
        validate_checkpoint(ck, R, [N], nchunk, False)


def test_checkpoint_items_and_sizes_are_checked_on_the_host():
    import numpy as np

    from frictionqpotspringblock_amd.batch import validate_checkpoint

    R, N, nchunk = 2, 5, 100
    ck = _checkpoint(R, N, nchunk)
    del ck["st_end"]
    with pytest.raises(RuntimeError, match="st_end is missing"):
        validate_checkpoint(ck, R, [N], nchunk, False)
    ck = _checkpoint(R, N, nchunk)
    with pytest.raises(RuntimeError, match="y has"):
        validate_checkpoint(ck, R, [N], nchunk + 1, False)
    with pytest.raises(RuntimeError, match="f_thermal is missing"):
        validate_checkpoint(ck, R, [N], nchunk, True)
    ck["y"][3] = np.nan
    with pytest.raises(RuntimeError, match="not finite"):
        validate_checkpoint(ck, R, [N], nchunk, False)

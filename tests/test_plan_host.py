"""Plan model of the plain path (atrip_hip_plan_model), host only: the window
edge and doubles tile that atrip_hip_run takes for a shape.  The pinned picks
are the ones the kernel traces of the measured runs show (doubles_f64_kernel
<5, 5> at No=40 and No=60, doubles_tri_kernel<2, 5> at No=40 and <3, 4> at
No=60) and the measured choices on the served stream's large shapes."""
import pytest

from atrip_amd import hip

MIXED = ("split_fp32", "fp32", "bf16x3")


def test_headline_shapes():
    assert hip.plan_model(40, 400, "fp64") == (8, 5, 5)
    assert hip.plan_model(60, 600, "fp64") == (8, 5, 5)
    for form in MIXED:  # one 32x32 kernel family, one model
        assert hip.plan_model(40, 400, form) == (8, 2, 5)
        assert hip.plan_model(60, 600, form) == (8, 3, 4)


def test_served_shapes():
    # No=26 Nv=308: edge 6 (40.6 against 33.2 TFLOP/s with edge 8), tile 4x5;
    # No=31 Nv=270: 5x4; No=38 Nv=204: 4x5; No=20 Nv=183: 5x5
    assert hip.plan_model(26, 308, "fp64") == (6, 4, 5)
    assert hip.plan_model(31, 270, "fp64") == (8, 5, 4)
    assert hip.plan_model(38, 204, "fp64") == (8, 4, 5)
    assert hip.plan_model(20, 183, "fp64") == (8, 5, 5)


def test_forced_window_and_tile_ranges():
    for w in (1, 6, 12, 32):
        nb, fm, fn = hip.plan_model(40, 400, "fp64", window=w)
        assert nb == w and 2 <= fm <= 5 and 2 <= fn <= 5
    for no in (1, 3, 8, 17, 64):
        for nv in (1, 2, 50, 333):
            nb, fm, fn = hip.plan_model(no, nv, "bf16x3")
            assert nb in (6, 8, 12) and 2 <= fm <= 4 and 2 <= fn <= 5


@pytest.mark.parametrize("args", [(0, 10, "fp64", 0), (4, 0, "fp64", 0),
                                  (4, 10, "fp64", -1), (4, 10, "fp64", 33)])
def test_bad_shapes_are_errors(args):
    with pytest.raises(hip.HipError):
        hip.plan_model(*args)


def test_unknown_precision_is_an_error():
    L = hip.load()
    assert L.atrip_hip_plan_model(4, 10, 4, 0, None, None, None) != 0
    assert L.atrip_hip_plan_model(4, 10, -1, 0, None, None, None) != 0
    assert L.atrip_hip_plan_model(4, 10, 0, 0, None, None, None) == 0

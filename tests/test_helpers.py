"""The ofHelpers.py glue of the shim (scenes/ofHelpers.py:9-100) against
hand-worked values (no GPU: these are host functions).  The reference's vec4
components are single precision, so every stored component is the float32
of the Python expression."""
import numpy as np
import pytest

from ofblend_amd import manta as M


def test_init_res():
    g3, g4 = M.initRes(64, 1.5, 0.75, 1.)
    assert tuple(g3) == (64, 96, 64)
    assert tuple(g4) == (64.0, 96.0, 64.0, 48.0)
    g3, g4 = M.initRes(30, 1, 1.0 / 3.0, 1.)   # 10.000000 as float32
    assert tuple(g3) == (30, 30, 30) and g4[3] == float(np.float32(10.0))


def test_init_offsets_flof_settings():
    """flof.py:313-318,857: autoBorder 0.1, repeatStartDist = res * 0.1."""
    gs4d = M.initRes(64, 1, 1, 1.)[1]
    gsd = M.initRes(32, 1, 1, 1.)[1]
    lo, ls, do, ds, df = M.initOffsets(gs4d, gsd, 0.1, 6.4, 1.)
    f64 = 6.400000095367432            # float32(64 * 0.1)
    assert tuple(lo) == (f64, f64, f64, 12.800000190734863)  # f32(f64 + 6.4)
    assert tuple(ls) == (0.800000011920929,) * 4            # f32(1 - 0.2)
    assert tuple(do) == (0.0,) * 4 and tuple(ds) == (1.0, 1.0, 1.0, 1.0)
    assert tuple(df) == (2.0,) * 4
    _, ls, _, ds, df = M.initOffsets(gs4d, M.initRes(48, 1, 1, 1.)[1], 0.25, 0., 0.5)
    assert ls[3] == 0.25 and ds[3] == 0.5
    assert df[0] == float(np.float32(64.0 / 48.0))


def test_vec4_helpers_and_params():
    v = M.Vec4(1.9, 2.5, -3.7, 4.2)
    assert tuple(M.vec4ToVec3i(v)) == (1, 2, -3)
    assert tuple(M.vec4ToVec3(v)) == tuple(float(np.float32(c)) for c in (1.9, 2.5, -3.7))
    assert M.vec4intString(v) == "[1,2,-3,4]"
    argv = ["flof.py", "dataid0", "3", "mode", "1"]
    used = []
    assert M.getParam("dataid0", "0", used, argv) == "3"
    assert M.getParam("mode", "0", used, argv) == "1"
    assert M.getParam("missing", "x", used, argv) == "x"
    M.checkUnusedParam(used, argv)
    with pytest.raises(SystemExit):
        M.checkUnusedParam([0, 0, 0, 1, 1], argv)

"""The C-ABI library loads and exports every entry point include/mskg.h declares."""
import ctypes
import os
import re

from mulskips_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mskg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mskg_[a-z_0-9]+)\s*\(", src)))


def test_header_and_exports_agree():
    names = declared()
    assert len(names) >= 14
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names


def test_create_without_gpu_fails_cleanly():
    """mskg_create reports MSKG_EDEVICE (or succeeds where a GPU is present) -- never crashes."""
    import numpy as np
    from mulskips_amd.engine import Engine, EngineError
    t = dict(ptrans_e=np.zeros(12288), ptrans_d=np.ones(9), ptrans_abs=np.zeros(12), ptrans_des=np.zeros(12),
             ptrans_zig=1.0, n_crystal=1, n_cov=0, list_crystal=[14, 0, 0], list_cov=[0, 0, 0])
    try:
        e = Engine((48, 48, 48), t)
        e.close()
    except EngineError as err:
        assert err.code == _lib.MSKG_EDEVICE


def test_bad_config_rejected():
    import numpy as np
    from mulskips_amd.engine import Engine, EngineError
    t = dict(ptrans_e=np.zeros(12288), ptrans_d=np.ones(9), ptrans_abs=np.zeros(12), ptrans_des=np.zeros(12),
             ptrans_zig=1.0, n_crystal=1, n_cov=0, list_crystal=[14, 0, 0], list_cov=[0, 0, 0])
    try:
        Engine((50, 48, 48), t)
        raise AssertionError("box 50 accepted")
    except EngineError as err:
        assert err.code == _lib.MSKG_EINVAL

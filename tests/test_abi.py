"""The C-ABI library loads, exports every declared symbol and reports errors."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cfd_mi355x.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("cfd_create", "cfd_configure", "cfd_set_types", "cfd_init_equilibrium",
                 "cfd_set_populations", "cfd_set_front", "cfd_step", "cfd_forces",
                 "cfd_read_macros", "cfd_update_macros", "cfd_voxelize", "cfd_destroy",
                 "cfd_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from cfd__amd import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared_functions())


def test_version_string():
    from cfd__amd import _lib
    assert "gfx950" in _lib.version()


def test_struct_layout_version_matches_the_header():
    """The Python structs, the header's CFD_ABI_VERSION and the library agree,
    and the loader refuses a library of another layout version."""
    from cfd__amd import _lib
    m = re.search(r"#define CFD_ABI_VERSION (\d+)", HEADER.read_text())
    assert m and int(m.group(1)) == _lib.ABI_VERSION == _lib.lib.cfd_abi_version()
    assert C.sizeof(_lib.Tuning) == 6 * 4          # cfd_tuning with rotate
    assert C.sizeof(_lib.Params) == 12 * 8
    assert C.sizeof(_lib.Mlp) == 1378 * 4          # cfd_mlp with activation
    assert C.sizeof(_lib.Closure) == (1378 + 33) * 4
    hdr = HEADER.read_text()
    for nome, v in (("CFD_ACT_SILU", _lib.CFD_ACT_SILU), ("CFD_ACT_HARD_SILU", _lib.CFD_ACT_HARD_SILU)):
        assert re.search(rf"#define {nome} {v}\b", hdr), nome
    src = (ROOT / "cfd__amd" / "_lib.py").read_text()
    assert "abi != ABI_VERSION" in src
    addon = (ROOT / "cfd__amd" / "node" / "addon.cc").read_text()
    assert "cfd_abi_version() != CFD_ABI_VERSION" in addon


def test_argument_errors_without_device_work():
    """Argument validation happens before any device call."""
    from cfd__amd import _lib
    h = C.c_void_p()
    rc = _lib.lib.cfd_create(0, 4, 4, 0, C.byref(h))
    assert rc == -1 and "must be positive" in _lib.lib.cfd_last_error().decode()
    rc = _lib.lib.cfd_create(4, 4, 4, 7, C.byref(h))
    assert rc == -1 and "flags" in _lib.lib.cfd_last_error().decode()
    rc = _lib.lib.cfd_create(2048, 1024, 1024, 0, C.byref(h))
    assert rc == -1 and "2^30" in _lib.lib.cfd_last_error().decode()
    rc = _lib.lib.cfd_step(None, 1)
    assert rc == -1 and "null" in _lib.lib.cfd_last_error().decode()
    rc = _lib.lib.cfd_voxelize(None, 0, None, 0, 4, 4, 4, float("nan"), 1,
                               np.zeros(64, np.uint8).ctypes.data, None, None)
    assert rc == -1 and "finite" in _lib.lib.cfd_last_error().decode()


def _has_device():
    from cfd__amd import _lib
    h = C.c_void_p()
    rc = _lib.lib.cfd_create(4, 4, 4, 0, C.byref(h))
    if rc == 0:
        _lib.lib.cfd_destroy(h)
    return rc == 0


def test_no_device_raises_loudly():
    """Without a GPU the product path fails with an error; there is no CPU
    fallback behind the ABI."""
    from cfd__amd import _lib
    if _has_device():
        pytest.skip("a HIP device is visible")
    from cfd__amd.solver import SolverMI355X
    with pytest.raises(_lib.CFDError) as e:
        SolverMI355X.criar({"nx": 8, "ny": 8, "nz": 8})
    assert "device" in str(e.value).lower()


def test_batch_and_closure_argument_errors():
    """The batch and closure entry points reject bad arguments before any
    device call, with a message naming the entry point."""
    from cfd__amd import _lib
    L = _lib.lib
    out = C.c_void_p()
    assert L.cfd_batch_create(None, 0, C.byref(out)) == -1
    assert "cfd_batch_create" in L.cfd_last_error().decode()
    assert L.cfd_batch_run(None, 10, 0, None, None) == -1
    assert "cfd_batch_run" in L.cfd_last_error().decode()
    assert L.cfd_batch_max_grid(256, None, None) == -1
    assert "cfd_batch_max_grid" in L.cfd_last_error().decode()
    X = np.zeros(8 * 64, np.float32)
    y = np.zeros(64, np.float32)
    params = _lib.Mlp()
    cfg = _lib.TrainCfg(steps=10, batch=100, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1)
    rc = L.cfd_mlp_train(X.ctypes.data, y.ctypes.data, 64, C.byref(cfg), C.byref(params), None)
    assert rc == -1 and "multiple of 64" in L.cfd_last_error().decode()
    rc = L.cfd_mlp_infer(None, 0, None, None, None)
    assert rc == -1 and "cfd_mlp_infer" in L.cfd_last_error().decode()
    # an activation the library does not know is refused by every MLP entry point
    params.activation = 7
    cfg.batch = 64
    rc = L.cfd_mlp_train(X.ctypes.data, y.ctypes.data, 64, C.byref(cfg), C.byref(params), None)
    assert rc == -1 and "known activation" in L.cfd_last_error().decode()
    grad = _lib.Mlp()
    loss = C.c_double()
    rc = L.cfd_mlp_grad(X.ctypes.data, y.ctypes.data, 64, C.byref(params), C.byref(grad),
                        C.byref(loss))
    assert rc == -1 and "cfd_mlp_grad" in L.cfd_last_error().decode()
    cl = _lib.Closure()
    cl.mlp = params
    nu = np.zeros(64, np.float32)
    inside = np.zeros(64, np.uint8)
    rc = L.cfd_mlp_infer(X.ctypes.data, 64, C.byref(cl), nu.ctypes.data, inside.ctypes.data)
    assert rc == -1 and "cfd_mlp_infer" in L.cfd_last_error().decode()
    count = C.c_int64()
    assert L.cfd_closure_fallbacks(None, C.byref(count), 0) == -1

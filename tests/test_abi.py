"""The C-ABI library loads without a GPU and exports every function that
include/rr_fitness.h declares; calls that need a device fail cleanly."""
import ctypes
import os
import subprocess

import pytest

import capi


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = capi.header_functions()
    assert {"rr_create", "rr_destroy", "rr_fitness_batch", "rr_dataset_upload", "rr_predict"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_header_constants_consistent():
    D = capi.D
    assert D["RR_MAX_STACK"] == 8 and D["RR_MAX_COLS"] >= 64
    ops = [D[k] for k in ("RR_OP_LDV", "RR_OP_LDC", "RR_OP_PUSH", "RR_OP_SIN", "RR_OP_COS", "RR_OP_LN",
                          "RR_OP_EXP", "RR_OP_SQRT", "RR_OP_SQR")]
    assert len(set(ops)) == len(ops) and max(ops) < D["RR_OP_BIN"]
    assert D["RR_OP_BIN"] + 8 * D["RR_BOP_MAX"] + D["RR_M_CR"] < 256


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_no_device_fails_loudly():
    lib = capi.lib()
    p = ctypes.c_void_p()
    rc = lib.rr_create(0, ctypes.byref(p))
    assert rc != 0 and b"device" in lib.rr_last_error()
    assert lib.rr_fitness_batch(None, 0, None, 0, 0, None, None) != 0


def test_product_module_refuses_to_run_without_device():
    rrc = pytest.importorskip("rils_rols_cpp")
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    import numpy as np
    rr = rrc.rils_rols(False, 100, 10, 0.001, 50, 1.0, False, 0)
    with pytest.raises(RuntimeError, match="HIP device"):
        rr.fit(np.zeros(4), np.zeros(2), 2, 2)
    with pytest.raises(ValueError):
        rr.fit(np.zeros(5), np.zeros(2), 2, 2)     # size mismatch: the reference exits, this raises

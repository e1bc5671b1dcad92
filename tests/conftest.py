import glob
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcpim_gpu.so)")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    from cpim_amd.model import ModelDesc
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    desc = ModelDesc(z["dom_size"], z["con_arity"], z["con_scope"], z["bin_rel"], name=name)
    return desc, {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu():
    """The GPU and the native library must be there: a GPU test never skips."""
    from cpim_amd import _native as N
    lib = N.gpu_lib()
    import ctypes as C
    info = N.CDeviceInfo()
    code = lib.cpim_get_device_info(0, C.byref(info))
    if code != 0:
        pytest.fail("no GPU visible: " + lib.cpim_last_error().decode())
    return info

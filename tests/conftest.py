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


def large_table_model(seed, n=12, d=5, e=14, p=0.35,
                      tabs=((6, 4500), (6, 4500), (7, 5000), (8, 5000))):
    """Seeded random binary network plus positive tables of arity 6 to 8 with
    4500 to 5000 tuples each (two to three 4096-tuple chunks per table)."""
    from cpim_amd import workload
    from cpim_amd.model import from_tuples
    base = workload.random_binary(n, d, e, p, seed)
    scopes = [list(s) for s in base.scopes()]
    lists = list(base.tuple_lists())
    rng = np.random.default_rng(seed)
    for k, (ar, nt) in enumerate(tabs):
        scopes.append(sorted(rng.choice(n, ar, replace=False).tolist()))
        lists.append(workload.random_table(seed * 131 + k, ar, d, nt))
    return from_tuples(base.dom_size, scopes, lists)

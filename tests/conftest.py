import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device)")


@pytest.fixture(scope="session")
def adi():
    from phd_amd import chemdata as cd
    return cd.Adim()


@pytest.fixture(scope="session")
def thd(adi):
    from phd_amd import chemdata as cd
    return cd.build_thd(adi)


@pytest.fixture(scope="session")
def mech():
    from phd_amd import chemdata as cd
    return cd.read_mechanism()


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def host_bdf(tmp_path_factory, mech):
    """g++ build of the chemistry integrator source that the kernels compile
    (tests/host_bdf.cpp), set up with the bundled mechanism."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    from phd_amd.solver import mech_flat
    so = tmp_path_factory.mktemp("hostbdf") / "libhostbdf.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I",
                    os.path.join(ROOT, "phd_amd", "csrc"), "-o", str(so),
                    os.path.join(ROOT, "tests", "host_bdf.cpp")], check=True)
    L = C.CDLL(str(so))
    L.host_setup.argtypes = [C.c_void_p, C.c_long]
    L.host_psr.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                           C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    L.host_rhs.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    L.host_jac.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
    mf = mech_flat(mech)
    assert L.host_setup(mf.ctypes.data, mf.size) == 0
    return L

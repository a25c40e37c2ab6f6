"""Shared fixtures: the product binding (loaded by path: the package directory
name carries a hyphen), the oracle (test infrastructure only), the compiled
reference when present, and the synthetic frame generators."""
import ctypes
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "topos-codec_amd")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liborc.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libtopos_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def load_pkg():
    name = "topos_codec_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; no GPU context is created)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def abi(pkg):
    return pkg.abi


@pytest.fixture(scope="session")
def syn(pkg):
    return pkg.load_synth()


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_LIB)
    lib.orc_crc32.restype = ctypes.c_uint32
    lib.orc_log2.restype = ctypes.c_double
    lib.orc_log2.argtypes = [ctypes.c_double]
    return lib


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    lib = ctypes.CDLL(REF_LIB)
    lib.tc_last_error.restype = ctypes.c_char_p
    lib.tc_frame_packet_bound.restype = ctypes.c_size_t
    return lib


@pytest.fixture(scope="session")
def codec(pkg):
    return pkg.DeviceCodec()

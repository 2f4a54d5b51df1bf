"""The C-ABI libraries load and export every entry point their headers
declare (no compute: no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"simple_ts.h": "libsimple_ts_hip.so", "simple_ts_host.h": "libsimple_ts_host.so",
           "sts_unet.h": "libsimple_ts_hip.so"}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sts_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports(header):
    names = declared(header)
    assert len(names) >= 6
    lib = ctypes.CDLL(os.path.join(ROOT, "simple-ts_amd", HEADERS[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{HEADERS[header]} lacks {missing}"


def test_driver_built():
    path = os.path.join(ROOT, "simple-ts_amd", "sts_gpu")
    assert os.access(path, os.X_OK)

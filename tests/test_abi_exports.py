"""The C-ABI libraries load and export every function include/gatt_dns.h and
include/gatt_flow_io.h declare (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gatt_dns.h")
FLOW_HEADER = os.path.join(ROOT, "include", "gatt_flow_io.h")


def declared(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(gatt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_runtime_exports():
    from gatt_dns_amd.runtime import EXPORTS
    assert sorted(EXPORTS) == declared()


def test_library_exports_every_declared_symbol():
    from gatt_dns_amd.runtime import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("libgatt_dns.so not built")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_flow_library_exports_every_declared_symbol():
    from gatt_dns_amd import flowio
    assert sorted(flowio.EXPORTS) == declared(FLOW_HEADER)
    if not os.path.exists(flowio.LIB_PATH):
        pytest.skip("libgatt_flowio.so not built")
    lib = ctypes.CDLL(flowio.LIB_PATH)
    missing = [n for n in declared(FLOW_HEADER) if not hasattr(lib, n)]
    assert not missing, missing

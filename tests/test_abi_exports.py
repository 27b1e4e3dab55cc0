"""The C-ABI library loads and exports every function include/gatt_dns.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gatt_dns.h")


def declared():
    text = open(HEADER).read()
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

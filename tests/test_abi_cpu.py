"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
symbol declared in include/*.h, its Python binding covers them, and the
host-side error paths report cleanly."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            for m in re.finditer(r"\b(prr_[a-z0-9_]+)\s*\(", src):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from photorealisticrender4610_amd import abi
    L = abi.lib()
    names = declared_symbols()
    assert len(names) >= 18
    for n in sorted(names):
        assert hasattr(L, n), n
        assert n in abi.SIGNATURES, "python binding misses " + n


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from photorealisticrender4610_amd import abi
    h = C.c_void_p()
    rc = abi.lib().prr_create(0, C.byref(h))
    assert rc != 0
    assert abi.lib().prr_last_error(None)

"""The C-ABI library loads on the CPU host and exports every declared symbol
(no compute: those calls need a GPU)."""
import ctypes
import os
import re

import pytest

from fasterhenry_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fname in os.listdir(os.path.join(ROOT, "include")):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fname)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(fh_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.fail(f"{native.LIB_PATH} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(native.SYMBOLS)


def test_resolve_and_version_need_no_gpu():
    lib = native.lib()
    assert lib.fh_resolve(0, 10_000) == 0
    assert lib.fh_resolve(0, 10_001) == 1
    assert lib.fh_resolve(1, 10 ** 9) == 0
    assert lib.fh_resolve(2, 1) == 1
    assert b"gfx950" in lib.fh_version()


def test_no_device_fails_loudly_on_the_cpu_host():
    """Without a HIP device a compute entry returns FH_E_HIP, never a CPU result."""
    import fasterhenry_amd as fa
    from fasterhenry_amd import fixtures

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    g, ports, d = fixtures.spiral_system()
    with pytest.raises(fa.HipError):
        fa.MeshSystem.assemble(g, ports, d)

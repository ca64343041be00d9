"""The product libraries load on a machine without a GPU and export every
entry point their headers declare (no device call is made here)."""
import ctypes
import os
import re
import subprocess

import support as S

PKG = os.path.join(S.ROOT, "robust-implicit-surface-networks_amd")


def _declared(header):
    text = open(os.path.join(S.ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rin_[a-z_0-9]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    names = _declared("rin.h")
    assert len(names) > 20
    lib = ctypes.CDLL(os.path.join(PKG, "librin.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_exports_reference_entry_points():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "librin_dropin.so")], capture_output=True,
                         text=True, check=True).stdout
    demangled = subprocess.run(["c++filt"], input=out, capture_output=True, text=True, check=True).stdout
    for fn in ("implicit_arrangement(", "material_interface(", "rin::set_device("):
        assert fn in demangled, fn
    # the drop-in links the C ABI library and the host graph stages
    ctypes.CDLL(os.path.join(PKG, "librin_dropin.so"))


def test_no_cpu_path_in_the_product():
    """The product libraries never link the oracle."""
    for so in ("librin.so", "librin_host.so", "librin_dropin.so", "librin_serve.so"):
        deps = subprocess.run(["ldd", os.path.join(PKG, so)], capture_output=True, text=True).stdout
        assert "oracle" not in deps, (so, deps)


def test_serving_library_exports_declared_symbols():
    text = open(os.path.join(S.ROOT, "include", "rin", "serve.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(rin_server_[a-z_0-9]+)\s*\(", text)))
    assert len(names) >= 9
    lib = ctypes.CDLL(os.path.join(PKG, "librin_serve.so"))
    assert not [n for n in names if not hasattr(lib, n)]


def test_serving_admission_rule():
    """rin_server_admits (include/rin/serve.h): the head job starts when
    nothing runs, or when its estimate fits beside the memory in use."""
    lib = ctypes.CDLL(os.path.join(PKG, "librin_serve.so"))
    f = lib.rin_server_admits
    f.argtypes = [ctypes.c_uint64] * 4
    f.restype = ctypes.c_int
    assert f(0, 0, 10**12, 1) == 1          # nothing runs: admit whatever the budget
    assert f(1, 60, 40, 100) == 1           # fits exactly
    assert f(1, 61, 40, 100) == 0           # one byte over
    assert f(3, 0, 100, 100) == 1
    assert f(2, 90, 0, 80) == 0             # already over budget

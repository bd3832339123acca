"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(m[a-z]+_[a-z_0-9]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_headers_parse():
    assert "mf_factorize" in declared("mfhip.h")
    assert "mfs_symbolic" in declared("mfsym.h")


def test_libmfsym_exports(ct):
    lib = ctypes.CDLL(os.path.join(ROOT, "cliquetrees.jl_amd", "libmfsym.so"))
    for name in declared("mfsym.h"):
        assert hasattr(lib, name), name


def test_libmfhip_exports(ct):
    path = os.path.join(ROOT, "cliquetrees.jl_amd", "libmfhip.so")
    lib = ctypes.CDLL(path)
    names = declared("mfhip.h")
    for name in names:
        assert hasattr(lib, name), name
    from cliquetrees_amd import _hipsig
    assert sorted(_hipsig.EXPORTS) == names


def test_oracle_loads(oracle):
    lib = oracle.lib()
    for name in ["or_chol_impl", "or_div_fwd", "or_div_bwd", "or_set_blas"]:
        assert hasattr(lib, name)

"""The C-ABI libraries load and export every function their headers declare."""
import ctypes
import os
import re

from openpstd_amd import _lib
from oracle import binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pstd_\w+|oracle_\w+)\s*\(", text, flags=re.M)))


def test_product_library_exports_header():
    names = _declared(os.path.join(ROOT, "include", "openpstd_amd.h"))
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_oracle_library_exports_header():
    names = _declared(os.path.join(ROOT, "oracle", "pstd_oracle.h"))
    lib = ctypes.CDLL(ob.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_product_does_not_link_oracle():
    """The product library carries no oracle code."""
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert not hasattr(lib, "oracle_create")
    src = ""
    for d, _, files in os.walk(os.path.join(ROOT, "openpstd_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".h", ".hip")):
                src += open(os.path.join(d, f)).read()
    assert "pstd_oracle" not in src and "from oracle" not in src and "import oracle" not in src

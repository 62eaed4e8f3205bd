"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)  # preprocessor lines (macros)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            if re.search(r"typedef[^;]*" + name, text):
                continue
            names.add(name)
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "odxBatchCreate" in names
    assert len(names) > 10


def test_library_exports_every_declared_symbol():
    from ode_amd import odx

    if not os.path.exists(odx.LIB_PATH):
        pytest.fail("libodx.so not built; run python -m ode_amd.build")
    lib = ctypes.CDLL(odx.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, "missing exports: %s" % missing


def test_snapshot_loader_rejects_missing_and_foreign_files(tmp_path):
    # host-side checks of odxBatchLoadSnapshot: no device work happens before them
    from ode_amd import odx
    L = odx.lib()
    assert not L.odxBatchLoadSnapshot(str(tmp_path / "absent.snap").encode())
    bad = tmp_path / "foreign.snap"
    bad.write_bytes(b"NOTASNAP" + bytes(200))
    assert not L.odxBatchLoadSnapshot(str(bad).encode())
    short = tmp_path / "short.snap"
    short.write_bytes(b"ODXSNAP1" + bytes(10))
    assert not L.odxBatchLoadSnapshot(str(short).encode())
    msgs = " ".join(m for _, m in odx.last_messages())
    assert "cannot open" in msgs and "not a complete snapshot" in msgs

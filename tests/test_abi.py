"""The C-ABI library loads and exports every function its headers declare
(no compute calls: runs without a GPU)."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{]*?)\)\s*;", text):
            name = m.group(1)
            pre = text[max(0, m.start() - 200):m.start()]
            if "typedef" in pre.split(";")[-1]:
                continue
            if name.startswith(("mc_", "mchip_")):
                names.add(name)
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "mchip_decode_chunks" in names
    assert "mc_archive_decode_chunk" in names or not os.path.exists(
        os.path.join(ROOT, "include", "matter_compressor_amd.h"))


def test_library_exports_every_declared_symbol():
    import mc_amd
    so = mc_amd.lib_path()
    if not os.path.exists(so):
        pytest.fail(f"{so} not built")
    L = C.CDLL(so)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    for n in declared_functions():
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_product_library_does_not_link_the_oracle():
    import mc_amd
    out = subprocess.run(["ldd", mc_amd.lib_path()], capture_output=True, text=True).stdout
    assert "oracle" not in out and "mcref" not in out
    nm = subprocess.run(["nm", "-D", mc_amd.lib_path()], capture_output=True, text=True).stdout
    assert "orc_" not in nm

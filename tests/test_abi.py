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
            if name.startswith(("mc_", "mchip_")) or name == "priors_load":
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


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "mc_archive.h")),
                    reason="reference sources absent (they exist only where the fixtures are made)")
def test_archive_struct_prefix_matches_reference_layout(tmp_path):
    """Reference modules compiled against the reference's private archive
    header (mc_cache.c and mc_volume.c read a->gen) must find the same field
    offsets in this library's struct."""
    probe = ("#include <stddef.h>\n#include <stdio.h>\n#include HDR\n"
             "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", offsetof(struct mc_archive, base),"
             " offsetof(struct mc_archive, cursor), offsetof(struct mc_archive, file_len),"
             " offsetof(struct mc_archive, quality), offsetof(struct mc_archive, gen));return 0;}\n")
    src = tmp_path / "probe.c"
    src.write_text(probe)
    outs = []
    for hdr, inc in (('"mc_host_internal.h"', ["-I", os.path.join(ROOT, "include"),
                                                 "-I", os.path.join(ROOT, "matter-compressor_amd", "host")]),
                     ('"mc_archive.h"', ["-I", REF_SRC, "-include", "matter_compressor.h"])):
        exe = tmp_path / f"probe{len(outs)}"
        subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", f"-DHDR={hdr}", *inc, str(src), "-o", str(exe)],
                       check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1], outs


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "mc_cache.c")),
                    reason="reference sources absent (they exist only where the fixtures are made)")
def test_reference_modules_link_against_this_library():
    """The reference's cache, render and stream modules link against
    libmc_amd.so in place of its codec/archive objects (no undefined symbol)."""
    import mc_amd
    assert os.path.exists(mc_amd.lib_path())
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mixlink"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libmcmix.so"))

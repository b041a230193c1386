"""The C-ABI libraries load and export every function their headers declare
(CPU only: symbol lookup, no compute calls)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# header -> library that implements it
HEADERS = {
    "kkmx_runtime.h": "kkmx_rt",
    "kkmx_paths.h": "kkmx_paths",
}


def declared_functions(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([a-z][a-z0-9_]*)\s*\(", text, re.M):
        if not m.group(1).startswith(("typedef", "struct")):
            names.add(m.group(1))
    return sorted(names)


@pytest.mark.parametrize("header,lib", sorted(HEADERS.items()))
def test_library_exports_every_declared_symbol(header, lib):
    import kokkos_amd as K

    names = declared_functions(ROOT / "include" / header)
    assert len(names) >= 5, names
    L = K.load(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_every_header_is_mapped():
    hs = {p.name for p in (ROOT / "include").glob("*.h")}
    # kkmx_tools.h only declares the Kokkos Tools C types and callback typedefs
    assert hs - {"kkmx_tools.h"} <= set(HEADERS), hs


def test_loader_fails_loudly_without_library(tmp_path, monkeypatch):
    import kokkos_amd as K

    monkeypatch.setattr(K, "LIBDIR", tmp_path)
    monkeypatch.setattr(K, "_libs", {})
    with pytest.raises(K.KkmxError):
        K.load("kkmx_paths")

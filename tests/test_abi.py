"""The C-ABI library loads and exports every symbol include/mkgpu.h declares
(no compute calls: runs without a GPU)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mkgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mk_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_symbols():
    syms = declared_symbols()
    assert "mk_map2tod_f32" in syms and "mk_t2m_plan_create" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from minkasi_amd.lib import mkgpu
    missing = [s for s in declared_symbols() if not hasattr(mkgpu.lib, s)]
    assert not missing, missing
    assert mkgpu.lib.mk_version() >= 1
    assert mkgpu.lib.mk_reduce_workspace_bytes() > 0


def test_bindings_cover_header():
    from minkasi_amd.lib import mkgpu
    missing = [s for s in declared_symbols() if s not in mkgpu._SIGS]
    assert not missing, missing

"""The HIP library loads on a host without a GPU and exports every entry point
that include/qmin.h declares (no device calls here)."""
import os
import re

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "qmin.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "qm_fire_minimize" in names and "qm_ldg_compute_forces" in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    import qmin
    if not os.path.exists(qmin.LIB_PATH):
        pytest.skip("libqmin.so not built (run __graft_entry__.build())")
    lib = qmin.library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(qmin.exported_symbols()) == declared()

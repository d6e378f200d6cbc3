"""The C-ABI library loads and exports every symbol include/sloam_gpu.h
declares (CPU only: no device calls)."""
import ctypes
import os
import re

import pytest

from sloam_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "sloam_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(sloam_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_functions():
    names = _declared_functions()
    assert "sloam_ctx_create" in names and "sloam_cyl_plan_exec" in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_header():
    assert set(_declared_functions()) <= set(_lib.PROTOTYPES)


def test_strerror_without_device():
    lib = _lib.load()
    assert lib.sloam_strerror(0) == b"ok"
    assert lib.sloam_strerror(-1) == b"invalid argument"


def test_invalid_args_rejected_without_device():
    lib = _lib.load()
    assert lib.sloam_ctx_create(0, None) == -1
    assert lib.sloam_sync(None) == -1
    assert lib.sloam_cyl_plan_exec(None, None, None, None, None, None, None, None) == -1

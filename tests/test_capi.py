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


# C struct name -> ctypes mirror in sloam_amd/_lib.py
STRUCTS = {
    "sloam_cyl_fit_cfg": "CylFitCfg",
    "sloam_pose_cfg": "PoseCfg",
    "sloam_proj_cfg": "ProjCfg",
    "sloam_ground_cfg": "GroundCfg",
    "sloam_plane_fit_cfg": "PlaneFitCfg",
    "sloam_line_fit_cfg": "LineFitCfg",
    "sloam_trellis_cfg": "TrellisCfg",
    "sloam_frontend_cfg": "FrontendCfg",
    "sloam_frontend_out": "FrontendOut",
    "sloam_frontend_host": "FrontendHost",
    "sloam_pn_train_cfg": "PnTrainCfg",
    "sloam_kf_params": "KfParams",
    "sloam_kf_inputs": "KfInputs",
    "sloam_kf_result_view": "KfResultView",
}


def test_ctypes_struct_layouts_match_c(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives
    the header's struct (compiled with gcc, no device code)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sloam_gpu.h"', "int main(void) {"]
    for c_name, py_name in STRUCTS.items():
        st = getattr(_lib, py_name)
        lines.append(f'  printf("{py_name} size %zu\\n", sizeof({c_name}));')
        for field, _ in st._fields_:
            lines.append(f'  printf("{py_name} {field} %zu\\n", offsetof({c_name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        py_name, field, value = line.split()
        st = getattr(_lib, py_name)
        if field == "size":
            assert ctypes.sizeof(st) == int(value), py_name
        else:
            assert getattr(st, field).offset == int(value), (py_name, field)

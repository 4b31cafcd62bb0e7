"""CPU checks of the C ABI: the library loads, exports every function that
include/hp2d.h declares, and the ctypes struct layouts match the header."""
import os
import re
import subprocess

import ctypes as C
import pytest

from hydropol2d_amd import _lib, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hp2d.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hp2d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include "hp2d.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(hp2d_run_config),'
                   ' sizeof(hp2d_run_inputs), sizeof(hp2d_run_report),'
                   ' offsetof(hp2d_run_config, n_lulc_classes), offsetof(hp2d_run_inputs, gw_Sy));'
                   'return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    assert got == [C.sizeof(abi.RunConfig), C.sizeof(abi.RunInputs), C.sizeof(abi.RunReport),
                   abi.RunConfig.n_lulc_classes.offset, abi.RunInputs.gw_Sy.offset]


def test_no_device_raises_instead_of_falling_back():
    if _lib.device_count() > 0:
        pytest.skip("a device is visible")
    from hydropol2d_amd import local_inertial
    import numpy as np
    with pytest.raises(_lib.HP2DError):
        local_inertial.Inertial_Solver(1, np.zeros((2, 2)), 1.0, np.ones((2, 2)), np.zeros((2, 2)),
                                       np.zeros((2, 2)))

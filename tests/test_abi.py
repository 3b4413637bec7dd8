"""CPU checks of the C-ABI boundary: the library loads, exports every function
declared in include/langmuir_gpu.h, its ctypes structures match the C layout,
and it reports the no-device case and bad arguments as status codes."""
import ctypes as C
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "langmuir_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from langmuir_amd import _abi
    L = _abi.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"{n} declared in langmuir_gpu.h but not exported"
    assert sorted(_abi.EXPORTED) == names


def test_struct_layout_matches_header():
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    from langmuir_amd._abi import LmgParams, LmgSimConfig, LmgSimState
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "langmuir_gpu.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(lmg_params), sizeof(lmg_sim_config), sizeof(lmg_sim_state));
  printf("%zu %zu %zu\n", offsetof(lmg_params, prefactor), offsetof(lmg_sim_config, rate),
         offsetof(lmg_sim_config, coupling));
  printf("%zu %zu %zu\n", offsetof(lmg_sim_state, rng_window), offsetof(lmg_sim_state, rng_consumed),
         offsetof(lmg_sim_state, current_step));
  return 0; }
"""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "l.c")
        open(p, "w").write(src)
        subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), "-o", os.path.join(d, "l"), p], check=True)
        out = subprocess.run([os.path.join(d, "l")], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    want = [C.sizeof(LmgParams), C.sizeof(LmgSimConfig), C.sizeof(LmgSimState),
            LmgParams.prefactor.offset, LmgSimConfig.rate.offset, LmgSimConfig.coupling.offset,
            LmgSimState.rng_window.offset, LmgSimState.rng_consumed.offset,
            LmgSimState.current_step.offset]
    assert vals == want


def test_create_validates_arguments_before_touching_a_device():
    from langmuir_amd import _abi
    L = _abi.lib()
    bad = _abi.LmgParams(grid_x=0, grid_y=4, grid_z=4, cutoff=5, hopping_range=1)
    h = C.c_void_p()
    assert L.lmg_create(C.byref(bad), C.byref(h)) == _abi.LMG_ERR_ARG
    assert b"grid" in L.lmg_last_error(None)
    bad = _abi.LmgParams(grid_x=4, grid_y=4, grid_z=4, cutoff=5, hopping_range=3)
    assert L.lmg_create(C.byref(bad), C.byref(h)) == _abi.LMG_ERR_ARG


def test_no_device_is_reported_not_faked():
    import langmuir_amd
    if langmuir_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(langmuir_amd.LmgError) as e:
        langmuir_amd.Context((8, 8, 8))
    assert e.value.code == -2


def test_null_context_calls_fail_cleanly():
    from langmuir_amd import _abi
    L = _abi.lib()
    assert L.lmg_sync(None) == _abi.LMG_ERR_ARG
    assert L.lmg_sim_step(None, 1) == _abi.LMG_ERR_ARG
    assert L.lmg_pair_potential(None, None, None, 0, None, 0, 0, None) == _abi.LMG_ERR_ARG


def test_product_does_not_reference_the_oracle():
    """The product package never imports, loads or calls the oracle."""
    pkg = os.path.join(ROOT, "langmuir_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in text and "lmoracle" not in text, f

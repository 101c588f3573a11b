"""The deconwolf method module against deconwolf's own headers.

deconwolf_amd/adapter/method_shb_hip.c is built inside a deconwolf source
tree; these tests compile it here with -fsyntax-only against the real
src/dw.h and src/fim.h (the only stand-in is tests/shim/fftw3.h, which
declares the two FFTW types those headers name), and check that the option
values of include/dwhip.h equal deconwolf's enums (src/dw.h:52-83), so a
caller can pass dw_opts fields straight through.  They also check that the
module calls none of deconwolf's CPU methods: unsupported jobs exit, they
never fall back to the reference.  Skipped where deconwolf's sources are
not present.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
ADAPTER = os.path.join(ROOT, "deconwolf_amd", "adapter")
TIFF_INC = "/opt/conda/include"

needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "dw.h")) or shutil.which("gcc") is None,
                               reason="deconwolf's sources or gcc not present")


def _incs():
    incs = ["-I" + os.path.join(ROOT, "tests", "shim"), "-I" + REF_SRC, "-I" + os.path.join(ROOT, "include"),
            "-I" + ADAPTER]
    if os.path.exists(os.path.join(TIFF_INC, "tiffio.h")):
        incs.append("-I" + TIFF_INC)
    return incs


@needs_ref
def test_method_module_compiles_against_dw_h():
    r = subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror"] + _incs() +
                       [os.path.join(ADAPTER, "method_shb_hip.c")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@needs_ref
def test_option_values_equal_deconwolf_enums(tmp_path):
    src = tmp_path / "enums.c"
    src.write_text("\n".join([
        '#include "dw.h"',
        '#include "dwhip.h"',
        "_Static_assert(DWHIP_METHOD_RL == DW_METHOD_RL, \"RL\");",
        "_Static_assert(DWHIP_METHOD_SHB == DW_METHOD_SHB, \"SHB\");",
        "_Static_assert(DWHIP_ITER_ABS == DW_ITER_ABS, \"abs\");",
        "_Static_assert(DWHIP_ITER_REL == DW_ITER_REL, \"rel\");",
        "_Static_assert(DWHIP_ITER_FIXED == DW_ITER_FIXED, \"fixed\");",
        "_Static_assert(DWHIP_METRIC_MSE == DW_METRIC_MSE, \"mse\");",
        "_Static_assert(DWHIP_METRIC_IDIV == DW_METRIC_IDIV, \"idiv\");",
        "_Static_assert(DWHIP_START_FLAT == DW_START_FLAT, \"flat\");",
        "_Static_assert(DWHIP_START_LP == DW_START_LP, \"lp\");",
        "_Static_assert(DWHIP_START_IDENTITY == DW_START_IDENTITY, \"identity\");",
        "",
    ]))
    r = subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-Werror"] + _incs() + [str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_method_module_never_runs_the_reference():
    """No call into deconwolf's own deconvolution methods from the module."""
    text = open(os.path.join(ADAPTER, "method_shb_hip.c")).read()
    code = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # comments name them as the contract
    for name in ("deconvolve_shb", "deconvolve_rl", "deconvolve_shb_cl", "deconvolve_shb_cl2", "iter_shb",
                 "iter_rl"):
        assert re.search(r"\b%s\s*\(" % name, code) is None, name
    assert '#include "method_shb.h"' not in code and '#include "method_rl.h"' not in code

"""The C ABI library loads and exports every entry point include/qiw.h declares (CPU)."""
import ctypes as C
import os
import re

import pytest

from qiw import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(REPO, "include", "qiw.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char \*)\s*\*?(qiw_\w+)\s*\(", hdr, re.M)))


def test_header_declarations_are_exported():
    lib = abi.load_library()
    decl = declared_symbols()
    assert len(decl) >= 17
    assert sorted(abi.EXPORTS) == decl
    for name in decl:
        assert hasattr(lib, name), name


def test_abi_version_and_no_gpu_error():
    lib = abi.load_library()
    assert lib.qiw_abi_version() == 6
    if os.environ.get("HIP_VISIBLE_DEVICES") == "" or not os.path.exists("/dev/kfd"):
        h = C.c_void_p()
        assert lib.qiw_create(C.byref(h), 0) == abi.QIW_E_NO_GPU


def test_struct_layout_matches_header():
    # offsets a C caller (or Julia's ccall with an isbits struct) relies on
    assert C.sizeof(abi.qiw_sbm) == 32
    assert abi.qiw_model.n_pairs.offset == 36 and abi.qiw_model.pair_op_i.offset == 40
    assert abi.qiw_run.N_samples.offset == 32 and C.sizeof(abi.qiw_run) == 56
    assert abi.qiw_run.rand_params.offset == 48
    assert abi.qiw_rand.seed.offset == 8 and C.sizeof(abi.qiw_rand) == 24
    assert C.sizeof(abi.qiw_stats) == 56 and abi.qiw_model.dlr_rank.offset == 104
    assert abi.qiw_warp_config.n_train.offset == 24 and C.sizeof(abi.qiw_warp_config) == 48
    assert abi.qiw_tuning.chunk_cost.offset == 8 and abi.qiw_tuning.best_ms.offset == 24
    assert C.sizeof(abi.qiw_tuning) == 32


def test_struct_layout_against_the_c_compiler(tmp_path):
    """The ctypes mirrors agree with what a C compiler makes of include/qiw.h."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include <stdio.h>
#include "qiw.h"
int main(void) {
    printf("%zu %zu %zu %zu\\n", sizeof(qiw_run), offsetof(qiw_run, rand_params), sizeof(qiw_rand),
           offsetof(qiw_rand, target_std));
    printf("%zu %zu %zu\\n", sizeof(qiw_model), sizeof(qiw_tuning), sizeof(qiw_stats));
    return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [C.sizeof(abi.qiw_run), abi.qiw_run.rand_params.offset, C.sizeof(abi.qiw_rand),
                                     abi.qiw_rand.target_std.offset, C.sizeof(abi.qiw_model),
                                     C.sizeof(abi.qiw_tuning), C.sizeof(abi.qiw_stats)]

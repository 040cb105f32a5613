"""The C-ABI library loads and exports every symbol include/satt.h declares
(no compute calls: CPU only)."""
import ctypes
import os

import satt


def test_header_declares_the_drop_in_entry_points():
    names = satt.header_symbols()
    for n in ("satt_fwd_lonlat", "satt_adj_lonlat", "satt_mask", "satt_interp2",
              "satt_forward_simulate", "satt_plan_create", "satt_plan_run",
              "satt_plan_results", "satt_plan_destroy"):
        assert n in names


def test_library_exports_every_header_symbol():
    assert os.path.exists(satt.LIB_PATH), "libsatt.so not built"
    lib = ctypes.CDLL(satt.LIB_PATH)
    missing = [n for n in satt.header_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string_needs_no_device():
    lib = satt.load()
    assert b"gfx950" in lib.satt_version()


def test_struct_layouts_match_header_sizes():
    # satt_batch: 4 ints, 6 pointers, int, 4 pointers... -> padded like C
    assert ctypes.sizeof(satt.Batch) == 4 * 4 + 6 * 8 + 8 + 4 * 8 + 4 * 8 + 8 + 8
    assert ctypes.sizeof(satt.Result) == 9 * 8
    # satt_plan_stats: 2 ints, long long, 3 doubles, 4 long longs, 6 doubles
    assert ctypes.sizeof(satt.PlanStats) == 8 + 8 + 3 * 8 + 4 * 8 + 6 * 8


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof of the ABI structs as gcc lays them out equals the ctypes mirror."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "satt.h"\nint main(void){printf('
                   '"%zu %zu %zu %zu %zu %zu %zu %zu", sizeof(satt_batch), sizeof(satt_result),'
                   ' sizeof(satt_plan_stats), sizeof(satt_server_config), sizeof(satt_request),'
                   ' sizeof(satt_response), sizeof(satt_server_stats), sizeof(satt_tune_result));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run([cc, "-I", os.path.dirname(satt.HEADER), str(src), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                            check=True).stdout.split()]
    assert sizes == [ctypes.sizeof(satt.Batch), ctypes.sizeof(satt.Result),
                     ctypes.sizeof(satt.PlanStats), ctypes.sizeof(satt.ServerConfig),
                     ctypes.sizeof(satt.Request), ctypes.sizeof(satt.Response),
                     ctypes.sizeof(satt.ServerStats), ctypes.sizeof(satt.TuneResult)]

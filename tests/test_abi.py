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
    assert ctypes.sizeof(satt.Result) == 8 * 8

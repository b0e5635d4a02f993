"""LDS layout of every fused-kernel instance (host-side check, no device
calls): for each grid size, rows per lane, waves per SIMD, launch mode and
coefficient type whose instance exists, every region a lane group addresses
in its LDS slice lies inside that slice, and the blocks a CU holds fit its
160 KiB of LDS (ck_fused_lds_check evaluates the kernel's own compile-time
layout functions)."""
import ctypes as C

import pytest

import croma_hip as ch


@pytest.mark.parametrize("n", [64, 128, 256, 512])
def test_every_instance_lds_in_bounds(n):
    L = ch.lib()
    L.ck_fused_lds_check.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_long)]
    checked = 0
    for r in (2, 4, 8, 16):
        if not L.ck_fused_supported(n, r):
            continue
        for occ in (1, 2):
            for mode in range(4):
                for tb in (8, 4):
                    if tb == 4 and not L.ck_fused_f32_supported(n, r, occ, mode):
                        continue
                    info = (C.c_long * 7)()
                    rc = L.ck_fused_lds_check(n, r, occ, mode, tb, info)
                    assert rc == 0, (n, r, occ, mode, tb, list(info))
                    assert info[6] <= 160 * 1024
                    checked += 1
    assert checked >= 16

"""CPU: the C-ABI library loads and exports every entry point that
include/triblock_hip.h declares (no device work is done here)."""
import pytest
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "triblock_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tbs_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("tbs_lineLU", "tbs_triblock_solveDU", "tbs_triblock_lineRes", "tbs_copyAtoB",
                 "tbs_addAtoB", "tbs_copyAtoBjac", "tbs_assembleTriBlocks", "tbs_plan_solve",
                 "tbs_batch_solve"):
        assert must in names


def test_library_exports_every_declared_symbol(tbs):
    from triblocksolve_amd import _capi
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (tbs_[A-Za-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # the ctypes table covers the same set
    assert set(_capi.SIGNATURES) == set(declared())


def test_host_only_calls_work_without_gpu(tbs):
    lib = tbs.load()
    assert lib.tbs_version().startswith(b"triblocksolve")
    import ctypes
    buf = (ctypes.c_int * 32)()
    n = lib.tbs_supported_nvar(buf, 32)
    assert 9 in list(buf)[:n]
    n2 = lib.tbs_plan_supported_nvar(buf, 32)
    assert 9 in list(buf)[:n2]


def test_plan_footprint_is_host_side(tbs):
    m = tbs.structured_mesh(1200, 1280)
    fp64 = tbs.Plan.footprint(m, tbs.TBS_FP64)
    fp32 = tbs.Plan.footprint(m, tbs.TBS_FP32)
    assert 8e9 < fp64 < 20e9 and fp32 < fp64


def test_plan_refuses_lines_past_descriptor_range(tbs):
    """A line's factor windows are addressed through a 32-bit buffer range
    whose out-of-range base is 2^31 bytes: at nvar 16 in FP64 one window holds
    4 line steps in 8 KiB, so a line of about 1.05M elements is the longest a
    plan takes.  Longer lines are refused up front (TBS_ERR_UNSUPPORTED), on
    the host, rather than addressed with a wrapped range."""
    ok = tbs.structured_mesh(1, 1_000_000, nvar=16)
    assert tbs.Plan.footprint(ok, tbs.TBS_FP64) > 0
    long_line = tbs.structured_mesh(1, 1_100_000, nvar=16)
    with pytest.raises(tbs.TbsError) as ei:
        tbs.Plan.footprint(long_line, tbs.TBS_FP64)
    from triblocksolve_amd._capi import TBS_ERR_UNSUPPORTED
    assert ei.value.code == TBS_ERR_UNSUPPORTED and "too long" in str(ei.value)
    # FP32 factors halve the window: the same line fits
    assert tbs.Plan.footprint(long_line, tbs.TBS_FP32) > 0


def test_variant_names(tbs):
    """Thomas variant codes (100 * prefetch windows + ring cycles) name the
    compiled template instance the autotuner reports."""
    assert tbs.variant_name(302) == "pf3_k2"
    assert tbs.variant_name(1001) == "pf10_k1"
    assert tbs.variant_name(10302) == "pf3_k2_l2"   # two lines per workgroup

"""The C-ABI library loads without a GPU and exports every symbol the header declares."""

import ctypes
import re
from pathlib import Path

import pytest

from maomi_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "maomi_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(maomi_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "maomi_persistent_sites" in names and len(names) >= 5


def test_library_exports_every_declared_symbol():
    if not _native.library_path().exists():
        pytest.fail("libmaomi_amd.so is not built; run __graft_entry__.build()")
    handle = ctypes.CDLL(str(_native.library_path()))
    missing = [n for n in declared_functions() if not hasattr(handle, n)]
    assert not missing, missing
    # every binding in the Python wrapper is declared in the header
    assert set(_native.exported_names()) <= set(declared_functions())


def test_error_reporting_without_gpu_work():
    lib = _native.lib()
    assert lib.maomi_version().startswith(b"maomi_amd")
    rc = lib.maomi_persistent_sites(None, None)
    assert rc == -1
    assert b"null descriptor" in lib.maomi_last_error()
    with pytest.raises(RuntimeError, match="null descriptor"):
        _native.check(rc, "maomi_persistent_sites")


def test_sites_descriptor_layout_matches_header():
    # 21 pointers + 12 int64 scalars + 3 x 5 int64 class tables, no padding; the complex
    # screen pointer, the work counter and the audit are the last members
    assert ctypes.sizeof(_native.SitesDesc) == 8 * (21 + 12 + 15)
    assert _native.SitesDesc.audit.offset == 8 * (20 + 12 + 15)
    assert ctypes.sizeof(_native.SitesAudit) == 8 * 8
    assert _native.SitesDesc.screen.offset == 8 * (18 + 12 + 15)
    assert _native.SitesDesc.queue.offset == 8 * (19 + 12 + 15)
    assert ctypes.sizeof(_native.SitesChoice) == 40
    assert _native.SitesChoice.ms.offset == 32


def test_sites_launch_and_plan_reject_bad_arguments_before_any_gpu_work():
    lib = _native.lib()
    d = _native.SitesDesc()
    one = 8  # non-null addresses: every call below returns before touching them
    for name in ("bases", "base_off", "phase", "edge", "q1d", "twiddle", "site_x", "site_y",
                 "site_n", "site_s", "order", "mask", "mask_off", "energy", "workspace"):
        setattr(d, name, one)
    d.P, d.H, d.W, d.Z, d.grid, d.uniform_n = 4, 300, 300, 8, 4, 256
    d.scratch_stride = 128 * 128  # below 256^2
    assert lib.maomi_persistent_sites(d, None) == -1
    assert b"scratch_stride" in lib.maomi_last_error()
    d.scratch_stride = 256 * 256
    d.psf_out, d.psf_z0, d.psf_nz, d.psf_c0, d.psf_w = one, 0, 3, 200, 64  # 200 + 64 > 256
    assert lib.maomi_persistent_sites(d, None) == -1
    assert b"focal-plane output window" in lib.maomi_last_error()
    d.psf_c0, d.psf_z0, d.psf_nz = 0, 7, 3  # planes 7..9 of 0..8
    assert lib.maomi_persistent_sites(d, None) == -1
    d.psf_out = None
    d.uniform_n = 96
    assert lib.maomi_persistent_sites(d, None) == -2
    assert b"uniform_n must be" in lib.maomi_last_error()
    h = ctypes.c_void_p()
    assert lib.maomi_sites_plan_create(None, ctypes.byref(h)) == -1
    d.uniform_n, d.P = 0, 0
    assert lib.maomi_sites_plan_create(d, ctypes.byref(h)) == -1 and not h.value
    d.P, d.phase16, d.slice_scale = 4, one, one
    assert lib.maomi_sites_plan_create(d, ctypes.byref(h)) == -2
    assert b"FP32-phase" in lib.maomi_last_error()
    assert lib.maomi_sites_plan_replay(None, None) == -1
    assert lib.maomi_sites_plan_capture(None, 4, 512, 1, 0) == -1
    assert b"schedule 0 or 1" in lib.maomi_last_error()
    d.phase16 = d.slice_scale = None
    d.H, d.W = 40000, 20000  # 8-byte screens of one slice: 6.4e9 bytes, beyond 32-bit ranges
    assert lib.maomi_persistent_sites(d, None) == -1
    assert b"exceeds 4 GiB" in lib.maomi_last_error()
    d.H, d.W = 300, 300
    d.P = (1 << 30) + 1  # site indices and the work counter would come near int32 overflow
    assert lib.maomi_persistent_sites(d, None) == -1
    assert b"bad sizes" in lib.maomi_last_error()
    assert lib.maomi_sites_plan_destroy(None) == 0


def test_pupil_params_layout_and_rejection():
    # 4 int32 + 8 float, then two doubles at offset 48
    assert ctypes.sizeof(_native.PupilParams) == 64
    assert _native.PupilParams.scale_x.offset == 48
    p = _native.PupilParams()
    p.kind = 2  # a pair beam needs the column phase
    rc = _native.lib().maomi_pupil_field(None, 8, 8, None, None, None, p, None)
    assert rc == -1 and b"maomi_pupil_field: bad arguments" in _native.lib().maomi_last_error()


def test_library_binds_to_torchs_hip_runtime():
    """Loading the library through the wrapper in a fresh process maps one HIP runtime: torch's
    libamdhip64.so.7 also serves the library's dependency (a second runtime would launch the
    library's kernels outside torch's device context)."""
    import subprocess
    import sys
    code = ("from maomi_amd import _native; _native.lib(); "
            "print(sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=str(HEADER.parents[1])).stdout
    runtimes = eval(out.strip().splitlines()[-1])
    assert len(runtimes) == 1, runtimes


def test_v4_finish_rejects_bad_sizes_before_any_gpu_work():
    lib = _native.lib()
    assert lib.maomi_v4_finish_f64_workspace(3, 1) == -1
    assert lib.maomi_v4_finish_f64_workspace(10, 14) == -1
    assert lib.maomi_v4_finish_f64_workspace(10, 2) == (7 * 5 + 6) * 8
    one = 8
    assert lib.maomi_v4_finish_f64(one, 10, 10, one, one, one, 14, one, one, None) == -1
    assert b"1 <= C <= 13" in lib.maomi_last_error()
    assert lib.maomi_v4_finish_f64(one, 9, 10, one, one, one, 2, one, one, None) == -1

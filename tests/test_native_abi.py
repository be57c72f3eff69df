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
    # 19 pointers + 12 int64 scalars + 3 x 5 int64 class tables, no padding; the complex
    # screen pointer is the last member
    assert ctypes.sizeof(_native.SitesDesc) == 8 * (19 + 12 + 15)
    assert _native.SitesDesc.screen.offset == 8 * (18 + 12 + 15)


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

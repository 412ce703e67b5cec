"""The C-ABI library loads without a GPU and exports every symbol include/pypty_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pypty_amd.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pty_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from pypty_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__

        __graft_entry__.build()
    return _lib.load()


def test_header_declares_the_expected_entry_points():
    from pypty_amd import _lib

    assert set(_declared()) == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_only_entry_points(lib):
    assert lib.pty_version() >= 100
    for n in (16, 32, 48, 64, 96, 128):
        assert lib.pty_supported_window(n) == 1
    for n in (0, 7, 100, 256):
        assert lib.pty_supported_window(n) == 0
    assert lib.pty_status_string(0) == b"ok"
    assert lib.pty_status_string(-4) == b"window outside the object"


def test_argument_errors_need_no_device(lib):
    from pypty_amd import _lib

    # null data / unsupported sizes are rejected before any HIP call
    assert lib.pty_fft2_c64(None, 4, 64, 0, None) == -1
    assert lib.pty_fft2_c64(None, 0, 100, 0, None) == -2
    assert lib.pty_fft2_c64(None, 0, 64, 9, None) == -1
    p = _lib.PtyProblem()
    p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = 100, 1, 1, 1
    size = ctypes.c_size_t(0)
    assert lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size)) == -2
    p.n, p.obj_ny, p.obj_nx, p.compute_batch = 64, 128, 128, 32
    assert lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size)) == 0
    assert size.value > 0
    b = _lib.PtyBuffers()
    assert lib.pty_loss_and_direction(ctypes.byref(p), ctypes.byref(b), None) == -1
    # plane transforms: every length plans (a prime factor above 31 as a direct stage, which needs a
    # second copy of the line in LDS), up to the LDS: 8192, or 4096 with such a factor
    fake = ctypes.c_void_p(256)
    assert lib.pty_fourier_clean_c64(fake, 1, 2 * 4099, 64, 0.5, 0.0, None) == -2
    assert lib.pty_fourier_clean_c64(fake, 1, 64, 2 * 4099, 0.5, 0.0, None) == -2
    assert lib.pty_plane_fft2_c64(fake, None, 1, 64, 16384, 0, None) == -2
    assert lib.pty_plane_fft2_c64(fake, None, 1, 160, 210, 0, None) == -5  # plans; no scratch given
    assert lib.pty_plane_fft2_c64(fake, None, 1, 94, 2 * 37, 0, None) == -5


def test_workspace_size_formula(lib):
    from pypty_amd import _lib

    p = _lib.PtyProblem()
    p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = 128, 10, 1, 5
    p.obj_ny = p.obj_nx = 512
    p.compute_batch, p.n_positions, p.obj_layout = 256, 4096, 1
    size = ctypes.c_size_t(0)
    assert lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size)) == 0
    n2 = 128 * 128 * 8
    stash = 256 * 5 * 10 * n2
    assert stash < size.value < 3 * stash


def test_checkpoint_segments_shrink_the_workspace(lib):
    """Thick sample, 256 px windows, 40 slices, 10 modes, 512 windows: keeping every 8th slice's wave
    (4 checkpoints + an 8-slice segment + the wave carried between segments) needs 13 wave slots per
    window and mode instead of 40."""
    from pypty_amd import _lib

    def ws(ck, step_obj=1):
        p = _lib.PtyProblem()
        p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = 256, 40, 1, 10
        p.obj_ny = p.obj_nx = 1536
        p.compute_batch, p.n_positions, p.obj_layout = 512, 16384, 1
        p.step_obj, p.checkpoint_slices = step_obj, ck
        size = ctypes.c_size_t(0)
        rc = lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size))
        return rc, size.value

    rc, full = ws(0)
    assert rc == 0
    rc, ck8 = ws(8)
    assert rc == 0
    # one wave slot for every window and mode of a launch; windows larger than 128 px cap the
    # launch so that one tile per workgroup fits in 12 GiB (all 512 windows here)
    windows = min(512, max(8, (12 << 30) // (40 * 256 * 256 * 8)))
    assert windows == 512
    slot = windows * 10 * 256 * 256 * 8
    assert full - ck8 == (40 - 13) * slot
    assert ws(40)[1] == full and ws(0, step_obj=0)[1] == ws(8, step_obj=0)[1]  # inactive cases keep every slice
    assert ws(-1)[0] == -1


def test_batch_lanes_add_one_set_of_per_batch_buffers(lib):
    """Every lane beyond the first adds exactly one more block of the per-batch buffers (the shared
    tables, probe partials and loss partials stay single); more than four lanes are rejected."""
    from pypty_amd import _lib

    def ws(lanes, n_slices=10):
        p = _lib.PtyProblem()
        p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = 128, n_slices, 1, 5
        p.obj_ny = p.obj_nx = 512
        p.compute_batch, p.n_positions, p.obj_layout, p.lanes = 256, 4096, 1, lanes
        size = ctypes.c_size_t(0)
        rc = lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size))
        return rc, size.value

    rc1, one = ws(1)
    rc2, two = ws(2)
    assert rc1 == rc2 == 0 and ws(0) == (0, one)
    stash = 256 * 5 * 10 * 128 * 128 * 8
    assert stash < two - one < one
    # a slice more adds a stash slot per window and mode and a band-limited window per window to
    # every lane (planar object layout: no object copies)
    slice_block = (256 * 5 + 256) * 128 * 128 * 8
    assert (ws(2, 11)[1] - two) == 2 * slice_block and (ws(1, 11)[1] - one) == slice_block
    assert ws(3)[1] - two == two - one and ws(4)[1] - ws(3)[1] == two - one
    assert ws(5)[0] == -1 and ws(-1)[0] == -1


def test_struct_layout_matches_header(tmp_path):
    """The ctypes mirrors of pty_problem and pty_buffers have the header's size and field offsets
    (a C program built with gcc from the header prints them)."""
    import shutil
    import subprocess

    from pypty_amd import _lib

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pypty_amd.h"', "int main(void) {"]
    for cname, cls in (("pty_problem", _lib.PtyProblem), ("pty_buffers", _lib.PtyBuffers)):
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for row in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = row.split()
        got[(cname, field)] = int(value)
    for cname, cls in (("pty_problem", _lib.PtyProblem), ("pty_buffers", _lib.PtyBuffers)):
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_lanes_are_limited_by_the_workspace_budget(lib):
    """Lanes whose per-batch buffers would take more than 128 GiB together are not planned: the thick
    sample at 512 windows keeping every slice (about 128 GB a lane) runs one lane whatever is asked,
    and with checkpoints every 8 slices (about 61 GB a lane) two."""
    from pypty_amd import _lib

    def ws(lanes, ck):
        p = _lib.PtyProblem()
        p.n, p.n_slices, p.n_obj_modes, p.n_probe_modes = 256, 40, 1, 10
        p.obj_ny = p.obj_nx = 1536
        p.compute_batch, p.n_positions, p.obj_layout, p.lanes = 512, 16384, 1, lanes
        p.step_obj, p.checkpoint_slices = 1, ck
        size = ctypes.c_size_t(0)
        assert lib.pty_workspace_bytes(ctypes.byref(p), ctypes.byref(size)) == 0
        return size.value

    assert ws(4, 0) == ws(1, 0)
    assert ws(4, 8) == ws(2, 8) > ws(1, 8)

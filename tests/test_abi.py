"""C-ABI checks that need no GPU: the library builds, loads, exports every
entry point include/pypo_hip.h declares, and fails loudly (never silently,
never through a CPU fallback) when no device is visible."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from pypo_amd import bind as B
from pypo_amd import geometry as G
from pypo_amd import structs as PS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pypo_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    # InterfaceCPU.h:24-61 FP64 symbol set, unchanged names
    for ref_name in ("propagateToGrid_JM", "propagateToGrid_EH", "propagateToGrid_JMEH",
                     "propagateToGrid_EHP", "propagateToGrid_scalar", "propagateToFarField",
                     "propagateRays", "generateGrid"):
        assert ref_name in names
    # InterfaceCPU.h:64-101 FP32 set and InterfaceCUDA.h:29-69
    for ref_name in ("propagateToGridf_JM", "propagateToGridf_EH", "propagateToGridf_JMEH",
                     "propagateToGridf_EHP", "propagateToGridf_scalar", "propagateToFarFieldf",
                     "callKernelf_JM", "callKernelf_EH", "callKernelf_JMEH", "callKernelf_EHP",
                     "callKernelf_FF", "callKernelf_scalar", "callRTKernel"):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(B.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_functions():
        assert re.search(rf"\bT {name}$", out, flags=re.M), name


def test_struct_layouts_match_the_reference_abi():
    # x86-64 SysV sizes of src/include/Structs.h:29-161
    assert ctypes.sizeof(PS.c2Bundle) == 12 * 8
    assert ctypes.sizeof(PS.c4Bundle) == 24 * 8
    assert ctypes.sizeof(PS.c2rBundle) == 15 * 8
    assert ctypes.sizeof(PS.reflcontainer) == 8 + 7 * 8
    assert ctypes.sizeof(PS.cframe) == 8 + 6 * 8
    assert PS.reflparams.flip.offset == 32 and PS.reflparams.gmode.offset == 36
    assert PS.reflparams.transf.offset == 72 and PS.reflparams.rms_seed.offset == 88
    assert ctypes.sizeof(PS.reflparams) == 96


def test_version_and_device_count_callable():
    lib = B.loadHIPlib()
    assert b"gfx950" in lib.pypo_hip_version()
    assert lib.pypo_hip_device_count() >= 0


def test_no_device_raises_instead_of_falling_back():
    lib = B.loadHIPlib()
    if lib.pypo_hip_device_count() > 0:
        pytest.skip("a device is visible here")
    src = G.plane({"gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    tgt = G.plane({"gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    from pypo_amd import types as T
    z = np.zeros((3, 3), complex)
    with pytest.raises(B.PyPOHIPError, match="no HIP device"):
        B.PyPO_HIPd(src, tgt, {"mode": "EH", "k": 1.0, "s_current": T.currents(z, z, z, z, z, z)})


def test_argument_errors_are_reported_not_fatal():
    lib = B.loadHIPlib()
    buf = ctypes.create_string_buffer(256)
    lib.pypo_hip_last_error(buf, 256)           # clear
    lib.pypo_hip_set_precision(7)
    assert lib.pypo_hip_last_error(buf, 256) != 0
    assert b"precision" in buf.value
    assert lib.pypo_hip_last_error(buf, 256) == 0    # cleared after reading
    lib.pypo_hip_set_rt_max_iter(0)
    assert lib.pypo_hip_last_error(buf, 256) != 0


def test_header_and_ctypes_layouts_agree(tmp_path):
    """sizeof / offsetof of every struct in include/pypo_hip.h, compiled with
    g++, against the ctypes twins (FP64 and the FP32 structs of Structs.h:216-362)."""
    structs = ["arrC1", "c2Bundle", "c4Bundle", "c2rBundle", "reflparams", "reflcontainer", "cframe",
               "arrC1f", "c2Bundlef", "c4Bundlef", "c2rBundlef", "reflparamsf", "reflcontainerf", "cframef",
               "aperparams"]
    lines = ["#include <cstdio>", "#include <cstddef>", f'#include "{HEADER}"', "int main() {"]
    for s in structs:
        lines.append(f'  printf("{s} SIZEOF %zu\\n", sizeof({s}));')
        for name, _ in getattr(PS, s)._fields_:
            lines.append(f'  printf("{s} {name} %zu\\n", offsetof({s}, {name}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++11", str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}
    for s in structs:
        st = getattr(PS, s)
        assert got[(s, "SIZEOF")] == ctypes.sizeof(st), s
        for name, _ in st._fields_:
            assert got[(s, name)] == getattr(st, name).offset, (s, name)
    # FP32 reflparams as the reference's ctypes lays it out (Structs.py:293-310)
    assert PS.reflparamsf.flip.offset == 32 and PS.reflparamsf.ecc_uv.offset == 48
    assert PS.reflparamsf.transf.offset == 64 and ctypes.sizeof(PS.reflparamsf) == 80


def test_new_entry_points_report_errors_without_a_device():
    """The kernel-selection hook rejects unknown kinds, and every entry point
    added for the figures of merit and the FP32 interface reports an error
    (here: no device) instead of crashing or computing on the CPU."""
    from pypo_amd import bind_gpu as BG
    from pypo_amd import efficiencies as E
    lib = B.loadHIPlib()
    buf = ctypes.create_string_buffer(256)
    lib.pypo_hip_last_error(buf, 256)
    assert lib.pypo_hip_set_ff_kernel(5) != 0
    assert lib.pypo_hip_last_error(buf, 256) != 0
    assert lib.pypo_hip_set_ff_kernel(1) == 0 and lib.pypo_hip_set_ff_kernel(0) == 0
    with pytest.raises(ValueError):
        B.set_ff_kernel("tensor")
    with pytest.raises(E.ApertureError):
        E.check_aperDict({"shape": "ellipse", "outer": [1, 1]})        # no "inner"
    with pytest.raises(E.ApertureError):
        E.aper_struct({"shape": "hexagon", "outer": [1, 1], "inner": [0, 0]})
    if lib.pypo_hip_device_count() > 0:
        pytest.skip("a device is visible here")
    E._declare(lib)
    out = (ctypes.c_double * 8)()
    d = ctypes.POINTER(ctypes.c_double)()
    c = PS.reflcontainer()
    assert lib.pypo_hip_aperture_sums_dev(ctypes.byref(c), d, d, E.aper_struct(None), out, None) != 0
    assert lib.pypo_hip_power_sums_dev(d, d, d, d, 3, out, None) != 0
    f = PS.cframe()
    assert lib.pypo_hip_frame_stats_dev(ctypes.byref(f), out, None) != 0
    lib.pypo_hip_last_error(buf, 256)
    # the FP32 entry points return void and record the failure
    BG.loadGPUlib()
    src = G.plane({"gmode": "xy", "lims_x": [-1, 1], "lims_y": [-1, 1], "gridsize": [3, 3]})
    from pypo_amd import types as T
    z = np.zeros((3, 3), complex)
    with pytest.raises(B.PyPOHIPError):
        BG.PyPO_GPUf(src, src, {"mode": "EH", "k": 1.0, "s_current": T.currents(z, z, z, z, z, z)})
    fr = T.frame(2, np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2))
    with pytest.raises(B.PyPOHIPError):
        BG.RT_GPUf({"fr_in": fr, "t_name": src, "tol": 1e-6, "t0": 1.0})
    with pytest.raises(B.PyPOHIPError):
        BG.PyPO_GPUf(src, src, {"mode": "FF", "k": 1.0, "s_current": T.currents(z, z, z, z, z, z)},
                     interface="cpu")
    assert lib.pypo_hip_stream_ordered_scratch() < 0
    assert lib.pypo_hip_scratch_bytes(None) < 0
    assert lib.pypo_hip_release_scratch(None) != 0
    lib.pypo_hip_last_error(buf, 256)

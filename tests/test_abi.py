"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/wrach.h declares, its struct layouts match the Python mirror, and
argument validation that happens before any device call returns the
documented error codes."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from tests.conftest import ROOT
from wrach_amd import _lib

HEADER = os.path.join(ROOT, "include", "wrach.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wrach_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_python_export_list():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_header_constants_match_python():
    text = open(HEADER).read()
    import re
    flag = int(re.search(r"#define WRACH_FLAG_HOST_PACK (0x[0-9a-fA-F]+)u", text).group(1), 16)
    assert flag == _lib.WRACH_FLAG_HOST_PACK
    kernels = int(re.search(r"#define WRACH_NUM_FRAME_KERNELS (\d+)", text).group(1))
    assert kernels == _lib.NUM_FRAME_KERNELS == len(_lib.FRAME_KERNEL_NAMES)
    assert int(re.search(r"#define WRACH_SOLVER_QUIET (0x[0-9a-fA-F]+)u", text).group(1), 16) == _lib.WRACH_SOLVER_QUIET
    for name in ("WRACH_SOLVER_R", "WRACH_SOLVER_S", "WRACH_MAX_PASSES", "WRACH_STENCIL_BUDGET"):
        value = int(re.search(rf"#define {name} (\d+)u", text).group(1))
        assert value == getattr(_lib, name), name


def test_library_exports_every_declared_symbol(wrach_lib_path):
    lib = ctypes.CDLL(wrach_lib_path)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", wrach_lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wrach_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_ctypes():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "wrach.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(wrach_config),
        offsetof(wrach_config, cell_size), offsetof(wrach_config, boundaries_as_dimensions),
        offsetof(wrach_config, gravity), offsetof(wrach_config, dt),
        offsetof(wrach_config, flags), offsetof(wrach_config, device),
        sizeof(wrach_world_settings), sizeof(wrach_device_view), sizeof(wrach_tuning),
        sizeof(wrach_solver_report), offsetof(wrach_solver_report, overlap_sum_before),
        offsetof(wrach_solver_report, pairs_after));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                              check=True).stdout.split()]
    C = _lib.WrachConfigC
    want = [ctypes.sizeof(C), C.cell_size.offset, C.boundaries_as_dimensions.offset,
            C.gravity.offset, C.dt.offset, C.flags.offset, C.device.offset,
            ctypes.sizeof(_lib.WorldSettingsC), ctypes.sizeof(_lib.DeviceViewC),
            ctypes.sizeof(_lib.TuningC), ctypes.sizeof(_lib.SolverReportC),
            _lib.SolverReportC.overlap_sum_before.offset, _lib.SolverReportC.pairs_after.offset]
    assert got == want
    assert got[7] == 32  # ShaderWorldSettings is 32 bytes (config_shader.rs:15-29)


def test_config_default_matches_reference(wrach_lib_path):
    L = _lib.load(wrach_lib_path)
    c = _lib.WrachConfigC()
    L.wrach_config_default(ctypes.byref(c))
    # config_app.rs:26-35: (480, 352), boundaries false, SPATIAL_BIN_CELL_SIZE = 3
    assert (c.width, c.height, c.cell_size, c.boundaries_as_dimensions) == (480, 352, 3, 0)
    assert (c.gravity[0], c.gravity[1], c.dt, c.flags) == (0.0, 0.0, 1.0, 0)


def test_argument_errors_before_any_device_call(wrach_lib_path):
    L = _lib.load(wrach_lib_path)
    h = ctypes.c_void_p()
    assert L.wrach_create(None, ctypes.byref(h)) == _lib.WRACH_E_INVALID
    assert b"null" in L.wrach_last_error()
    c = _lib.WrachConfigC()
    L.wrach_config_default(ctypes.byref(c))
    c.cell_size = 0
    assert L.wrach_create(ctypes.byref(c), ctypes.byref(h)) == _lib.WRACH_E_INVALID
    assert b"cell_size" in L.wrach_last_error()
    L.wrach_config_default(ctypes.byref(c))
    c.flags = 7
    assert L.wrach_create(ctypes.byref(c), ctypes.byref(h)) == _lib.WRACH_E_INVALID
    L.wrach_config_default(ctypes.byref(c))
    c.dt = float("nan")
    assert L.wrach_create(ctypes.byref(c), ctypes.byref(h)) == _lib.WRACH_E_INVALID
    # calls on a null context
    assert L.wrach_step(None, 1) == _lib.WRACH_E_INVALID
    assert L.wrach_sync(None) == _lib.WRACH_E_INVALID
    assert L.wrach_capacity(None, None, None) == _lib.WRACH_E_INVALID
    L.wrach_destroy(None)  # no-op


def test_package_fails_loudly_without_library(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib._lib_saved = _lib._lib
        try:
            _lib._lib = None
            _lib.load(str(tmp_path / "missing.so"))
        finally:
            _lib._lib = _lib._lib_saved


def test_product_does_not_reference_the_oracle():
    """The product path never imports, links or loads the oracle."""
    pkg = os.path.join(ROOT, "wrach_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f


RUST_LIB = os.path.join(ROOT, "wrach-hip-sys", "src", "lib.rs")


def test_rust_crate_binds_every_entry_point():
    """wrach-hip-sys (not compiled here: no cargo in this image) declares
    each function of the header once, with the same parameter count."""
    rust = open(RUST_LIB).read()
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name in declared_functions():
        m = re.search(rf"pub fn {name}\((.*?)\)", rust, flags=re.S)
        assert m, name
        c = re.search(rf"\b{name}\s*\((.*?)\)\s*;", text, flags=re.S)
        cparams = [p for p in c.group(1).split(",") if p.strip() and p.strip() != "void"]
        rparams = [p for p in m.group(1).split(",") if p.strip()]
        assert len(cparams) == len(rparams), name


def test_rust_crate_layouts_match_the_header():
    """The sizes and offsets the crate asserts at compile time are the
    header's, as a C compiler lays it out."""
    rust = open(RUST_LIB).read()
    sizes = dict((k, int(v)) for k, v in re.findall(r"size_of::<(\w+)>\(\) == (\d+)", rust))
    offsets = [(s, f, int(v)) for s, f, v in re.findall(r"offset_of!\((\w+), (\w+)\) == (\d+)", rust)]
    assert len(sizes) >= 6 and len(offsets) >= 4
    lines = [f'printf("{k} %zu\\n", sizeof({k}));' for k in sizes]
    lines += [f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));' for s, f, _ in offsets]
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"wrach.h\"\nint main(void) {\n" + \
          "\n".join(lines) + "\nreturn 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = dict(line.split() for line in subprocess.run([exe], capture_output=True, text=True,
                                                           check=True).stdout.splitlines())
    for k, v in sizes.items():
        assert int(got[k]) == v, k
    for s, f, v in offsets:
        assert int(got[f"{s}.{f}"]) == v, f"{s}.{f}"
    for name in ("WRACH_SOLVER_L", "WRACH_LEARN_PARAMS", "WRACH_LEARN_FEATURES", "WRACH_MAX_PASSES"):
        value = int(re.search(rf"#define {name} (\d+)u", open(HEADER).read()).group(1))
        assert int(re.search(rf"pub const {name}: u32 = (\d+);", rust).group(1)) == value, name
        assert getattr(_lib, name) == value, name

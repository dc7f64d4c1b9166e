"""CPU checks of the C-ABI library: it is built, loads, and exports every
symbol include/wgrid_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from wsclean_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "wgrid_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wgx?_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes signature table covers the same set
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_params_struct_layout():
    # the C struct: 4 size_t, 4 double, size_t, double, size_t, int (+pad)
    assert ctypes.sizeof(_lib.WgParams) == 4 * 8 + 4 * 8 + 8 + 8 + 8 + 8


def test_library_is_gfx950():
    so = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in so


def test_cpp_wrapper_compiles_and_links(tmp_path):
    # include/wgrid_hip.hpp (the WGriddingGridderBase-shaped C++ wrapper)
    # compiles against the C header and links against the library
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "wgrid_hip.hpp"\n'
        "#include <type_traits>\n"
        "static_assert(!std::is_copy_constructible<wgrid_hip::Gridder>::value, \"\");\n"
        "static_assert(!std::is_copy_constructible<wgrid_hip::Server>::value, \"\");\n"
        "int run(size_t n) {\n"
        "  wgrid_hip::Gridder g(n, n, n, n, 1e-5, 1e-5, 0.0, 0.0, 1, 1e-5);\n"
        "  g.InitializeInversion();\n"
        "  std::vector<float> img = g.RealImage();\n"
        "  wgrid_hip::Server s(2);\n"
        "  wg_request r{};\n"
        "  const uint64_t id = s.Submit(r);\n"
        "  const double t = s.Wait(id).ended;\n"
        "  return int(img.size() + g.ConstantMemoryUsage() + s.Stats()[0] + size_t(t));\n"
        "}\n"
        "int main(int argc, char**) { return argc > 5 ? run(64) : 0; }\n")
    exe = tmp_path / "use"
    lib_dir = os.path.dirname(_lib.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe), "-L", lib_dir, "-lwgrid_hip",
                        "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_request_struct_layout(tmp_path):
    # wg_request as the C compiler lays it out, against the ctypes mirror
    fields = [f for f, _ in _lib.WgRequest._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "wgrid_hip.h"\n'
                   "int main(void) {\n"
                   '  printf("%zu", sizeof(wg_request));\n' +
                   "".join('  printf(" %%zu", offsetof(wg_request, %s));\n' % f for f in fields) +
                   "  return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [ctypes.sizeof(_lib.WgRequest)] + [getattr(_lib.WgRequest, f).offset for f in fields]
    assert got == want

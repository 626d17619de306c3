"""The C-ABI library loads and exports every symbol include/remifa.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "remifa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(remifa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_helpers():
    syms = declared_symbols()
    for name in ["lu_nopiv_ll_f32", "lu_nopiv_ll_f32_fp32xtc32", "lu_nopiv_ll_f32_tc32xtc32",
                 "lu_nopiv_ll_f16", "lu_nopiv_ll_f16_tc16xtc16", "lu_nopiv_ll_f16_tc32xtc32",
                 "lu_nopiv_ll_f16_f16xtc32", "lu_nopiv_ll_f16_f16xtc32_fp32",
                 "lu_nopiv_ll_f16_f32xt32", "lu_nopiv_ll_f16_f16xtc16",
                 "lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32",
                 "lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32",
                 "lu_nopiv_panel", "convert", "gemmex"]:
        assert "remifa_" + name in syms


def test_library_exports_every_declared_symbol():
    import remifa_amd._lib as L
    lib = ctypes.CDLL(L.library_path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(declared_symbols()) <= set(L.SIGNATURES)


def test_python_package_fails_loudly_without_library(tmp_path, monkeypatch):
    import remifa_amd._lib as L
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "library_path", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        L.lib()


def test_cpp_layer_compiles_and_links(tmp_path):
    """include/remifa.hpp (the reference's names over the C ABI) compiles as
    host C++ and links against the library (nothing is run)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("no hipcc")
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "remifa.hpp"\n'
        "int main() {\n"
        "  remifa::Handle h;\n"
        "  float* a = nullptr; int* info = nullptr; __half* a16 = nullptr;\n"
        "  remifa::lu_nopiv_ll_f32(h, 0, 0, 256, a, 1, info);\n"
        "  remifa::lu_nopiv_ll_f32_tc32xtc32(h, 0, 0, 256, a, 1, info, a16, 1);\n"
        "  remifa::lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32(h, 0, 0, 256, a16, 1, info);\n"
        "  return 0;\n"
        "}\n")
    libdir = os.path.join(ROOT, "remifa_amd")
    exe = tmp_path / "caller"
    r = subprocess.run([hipcc, "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", libdir, "-lremifa_hip", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

"""CPU checks of the C ABI: the library loads and exports every entry point
that include/aldic.h declares; argument errors come back as codes."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "aldic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aldic_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    names = declared_symbols()
    assert len(names) >= 19
    for name in names:
        assert hasattr(lib, name), name


def test_bindings_cover_header(pkg):
    from importlib import import_module
    capi = import_module("aldic_amd._capi")
    assert set(declared_symbols()) == set(capi.SIGNATURES)


def test_null_context_is_an_error(pkg):
    capi = __import__("aldic_amd._capi", fromlist=["x"])
    lib = pkg.lib()
    assert lib.aldic_ctx_synchronize(None) == capi.EINVAL
    assert b"NULL" in lib.aldic_last_error()


def test_package_fails_loudly_without_library(tmp_path):
    import subprocess
    import sys
    code = (
        "import importlib.util, sys, shutil, os\n"
        f"src = {os.path.join(ROOT, '3d-stereo-aldic_amd')!r}\n"
        f"dst = {str(tmp_path / 'pkg')!r}\n"
        "shutil.copytree(src, dst, ignore=shutil.ignore_patterns('*.so', '_build'))\n"
        "spec = importlib.util.spec_from_file_location('aldic_amd', dst + '/__init__.py',"
        " submodule_search_locations=[dst])\n"
        "m = importlib.util.module_from_spec(spec); sys.modules['aldic_amd'] = m\n"
        "try:\n    spec.loader.exec_module(m)\nexcept ImportError as e:\n"
        "    print('IMPORTERROR', e); sys.exit(0)\n"
        "sys.exit(1)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode == 0 and "IMPORTERROR" in r.stdout

"""The C ABI library loads and exports what include/dsph.h declares (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT
from deorbit_amd import dsph

HEADER = os.path.join(ROOT, "include", "dsph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsph_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(dsph.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(dsph.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    lib.dsph_version.restype = C.c_char_p
    assert lib.dsph_version().startswith(b"dsph")


def test_struct_layouts_match_the_header(tmp_path):
    src = os.path.join(tmp_path, "sz.c")
    exe = os.path.join(tmp_path, "sz")
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include <stddef.h>\n#include "dsph.h"\n'
                "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(dsph_material),"
                " sizeof(dsph_params), sizeof(dsph_tuning), sizeof(dsph_diag), sizeof(dsph_profile),"
                " offsetof(dsph_params, g), offsetof(dsph_diag, steps), offsetof(dsph_material, Emod));"
                " return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(dsph.Material), C.sizeof(dsph.Params), C.sizeof(dsph.Tuning), C.sizeof(dsph.Diag),
            C.sizeof(dsph.Profile), dsph.Params.g.offset, dsph.Diag.steps.offset, dsph.Material.Emod.offset]
    assert got == want


def test_field_enum_matches_header():
    text = open(HEADER).read()
    body = re.search(r"typedef enum \{\s*(DSPH_F_POS.*?)\} dsph_field;", text, re.S).group(1)
    names = [n.strip().split("=")[0].strip() for n in body.replace("\n", " ").split(",") if n.strip()]
    assert names[-1] == "DSPH_NFIELDS"
    for k, name in enumerate(names[:-1]):
        short = name[len("DSPH_F_"):].lower()
        key = {"hh": "hh", "eps_work": "eps_work", "eps_act": "eps_act", "pf": "Pf", "rart": "Rart",
               "p": "P", "s": "S", "dsdt": "dSdt", "dsdt_old": "dSdt_old", "d": "D"}.get(short, short)
        assert dsph.FIELD[key] == k, name


def test_oracle_libraries_export_the_same_api():
    from oracle import pyoracle as po
    names = re.findall(r"\b(orc_[a-z_A-Z0-9]+)\s*\(", open(os.path.join(ROOT, "oracle", "sph_oracle.h")).read())
    libs = [po.ORACLE_SO] + ([po.REF_SO] if os.path.exists(po.REF_SO) else [])
    for path in libs:
        lib = C.CDLL(path)
        for name in set(names):
            assert hasattr(lib, name), (path, name)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(dsph, "_LIB", None)
    with pytest.raises(FileNotFoundError):
        dsph.load_library(os.path.join(tmp_path, "nope.so"))

"""The C-ABI library loads and exports every symbol include/gabe.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gabe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared("gabe.h")
    for must in ("gabe_create", "gabe_step", "gabe_energy", "gabe_spectra", "gabe_get_scalars"):
        assert must in names


@pytest.mark.parametrize("lib,header", [("gabe_amd/libgabe_amd.so", "gabe.h")])
def test_library_exports_every_declared_symbol(lib, header):
    path = os.path.join(ROOT, lib)
    if not os.path.exists(path):
        pytest.fail(f"{lib} not built (run __graft_entry__.build())")
    so = ctypes.CDLL(path)
    missing = [n for n in declared(header) if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    from gabe_amd import lattice
    assert sorted(lattice.EXPORTS) == declared("gabe.h")


def test_host_library_exports():
    path = os.path.join(ROOT, "gabe_amd", "libgabe_host.so")
    so = ctypes.CDLL(path)
    hdr = open(os.path.join(ROOT, "gabe_amd", "host", "gabe_host.h")).read()
    names = set(re.findall(r"\b(gabe_host_[a-z_]+)\s*\(", hdr))
    assert names and all(hasattr(so, n) for n in names)


def test_numbins_matches_reference_rule():
    from gabe_amd import lattice
    import math
    for N in (4, 16, 24, 64, 96, 256, 512):
        assert lattice.numbins(N) == int(math.sqrt(3.0) * N / 2 + 1)


def test_invalid_params_fail_loudly_without_gpu_work():
    from gabe_amd import lattice
    with pytest.raises(lattice.GabeError) as e:
        lattice.Lattice(N=15)
    assert e.value.code == lattice.GABE_ERR_ARG
    with pytest.raises(lattice.GabeError):
        lattice.Lattice(N=16, nflds=3)
    with pytest.raises(lattice.GabeError) as e:
        lattice.Lattice(N=16, stress_chunk=-2)
    assert e.value.code == lattice.GABE_ERR_ARG


def test_ctypes_mirror_matches_the_header_layout(tmp_path):
    """gabe_params and gabe_scalars as the C compiler lays them out from
    include/gabe.h: size and every field's offset equal the ctypes mirror's."""
    import subprocess
    from gabe_amd import lattice
    structs = {"gabe_params": lattice.GabeParams, "gabe_scalars": lattice.GabeScalars}
    # the mirror names two fields differently from the header
    cname = {"lam": "lambda"}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gabe.h"', "int main(void) {"]
    for st, py in structs.items():
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f, _ in py._fields_:
            c = cname.get(f, f)
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {c}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            st, f, v = line.split()
            got[(st, f)] = int(v)
    for st, py in structs.items():
        assert got[(st, "size")] == ctypes.sizeof(py), st
        for f, _ in py._fields_:
            assert got[(st, f)] == getattr(py, f).offset, (st, f)

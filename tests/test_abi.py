"""The C-ABI library loads and exports every symbol include/apophis_hydro.h declares,
and the numpy mirror of aph_block / aph_eos matches the C layout (no GPU calls)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from apophis_amd import _lib
from apophis_amd.mesh import BLOCK_DTYPE, EOS_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert sorted(_lib._SIGS) == names


def test_abi_version():
    assert _lib.lib().aph_abi_version() == 1


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    fields = [n for n in BLOCK_DTYPE.names]
    body = "".join(f'printf("%zu ", offsetof(aph_block, {f}));' for f in fields)
    efields = [n for n in EOS_DTYPE.names]
    ebody = "".join(f'printf("%zu ", offsetof(aph_eos, {f}));' for f in efields)
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"apophis_hydro.h\"\n"
        "int main(void){" + body + 'printf("%zu\\n", sizeof(aph_block));' + ebody +
        'printf("%zu\\n", sizeof(aph_eos)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    b_off = [int(v) for v in out[:len(fields)]]
    assert b_off == [BLOCK_DTYPE.fields[f][1] for f in fields]
    assert int(out[len(fields)]) == BLOCK_DTYPE.itemsize == 240
    e_off = [int(v) for v in out[len(fields) + 1:len(fields) + 1 + len(efields)]]
    assert e_off == [EOS_DTYPE.fields[f][1] for f in efields]
    assert int(out[-1]) == EOS_DTYPE.itemsize


def test_error_convention_without_device():
    """Null-argument calls return APH_E_INVALID without touching a device."""
    L = _lib.lib()
    assert L.aph_context_create(None, None) == -1
    assert L.aph_last_error(None) == b"null context"
    assert L.aph_blockset_ncells(None) == -1
    assert L.aph_cons_to_prim(None, None, 0, -1, None, None, None, None) == -1

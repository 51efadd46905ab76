"""The C-ABI library loads and exports every symbol include/giraffe_amd.h declares,
and the ctypes mirrors have the C layout (no GPU needed)."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

from giraffe_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = runtime.load()
    syms = runtime.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string():
    lib = runtime.load()
    assert lib.gf_version() >= 100
    assert isinstance(lib.gf_last_error(), bytes)


def test_null_arguments_are_rejected():
    lib = runtime.load()
    h = C.c_void_p()
    assert lib.gf_model_create(None, None, C.byref(h)) == abi.GF_E_ARG
    assert lib.gf_assemble(None, None) == abi.GF_E_ARG
    assert lib.gf_beam_eval(-1, None, None, None, None, None, None, None) == abi.GF_E_ARG


_PROBE = r'''
#include <stdio.h>
#include <stddef.h>
#include "giraffe_amd.h"
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("gf_model_desc size %zu\n", sizeof(gf_model_desc));
  printf("gf_static_params size %zu\n", sizeof(gf_static_params));
  printf("gf_static_report size %zu\n", sizeof(gf_static_report));
  printf("gf_options size %zu\n", sizeof(gf_options));
  FIELDS
  return 0;
}
'''


@pytest.mark.parametrize("ctype,cname", [(abi.GfModelDesc, "gf_model_desc"), (abi.GfStaticParams, "gf_static_params"),
                                         (abi.GfStaticReport, "gf_static_report"), (abi.GfOptions, "gf_options")])
def test_struct_layout_matches_header(ctype, cname):
    fields = "".join(f"P({cname}, {f})\n" for f, _ in ctype._fields_)
    src = _PROBE.replace("FIELDS", fields)
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode().split("\n")
    sizes = {l.split()[0]: int(l.split()[2]) for l in out if " size " in l}
    assert sizes[cname] == C.sizeof(ctype)
    offs = {l.split()[1]: int(l.split()[2]) for l in out if l.startswith(cname + " ")}
    for f, _ in ctype._fields_:
        assert offs[f] == getattr(ctype, f).offset, f

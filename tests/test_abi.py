"""The C ABI (include/*.h) as seen by a foreign caller: every declared entry
point is exported by the shared library, the parameter struct has the layout
the header gives it, and entry points that need no device behave as the
header says.  No call here touches a GPU."""
import ctypes as C
import os
import re

import pytest

import dnhip
import orc

DN_ERR_ARG = -1
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dn_[a-z0-9_]+|dnio_[a-z0-9_]+|dnfwi_[a-z0-9_]+)\s*\(", text)))


def _headers():
    return sorted(h for h in os.listdir(INCLUDE) if h.endswith(".h"))


def _lib_for(header):
    pkg = os.path.join(ROOT, "denise-black-edition_amd")
    return {"denise_hip.h": os.path.join(pkg, "libdnhip.so"),
            "denise_io.h": os.path.join(pkg, "libdnio.so"),
            "denise_fwi.h": os.path.join(pkg, "libdnio.so")}[header]


@pytest.mark.parametrize("header", _headers())
def test_every_declared_symbol_is_exported(header):
    names = _declared(header)
    assert names, header
    lib = C.CDLL(_lib_for(header))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    lib = dnhip.load()
    for n in _declared("denise_hip.h"):
        assert getattr(lib, n).restype is not None or n == "dn_ctx_destroy", n


def test_params_layout_matches_header():
    # 16 ints, 7 floats, fl[8], physics -- the order of dn_params in denise_hip.h
    assert C.sizeof(dnhip._CParams) == 16 * 4 + 7 * 4 + 8 * 4 + 4
    assert [f[0] for f in dnhip._CParams._fields_][:3] == ["nx", "ny", "nt"]
    assert dnhip._CParams._fields_[-1][0] == "physics"
    # the oracle takes the same struct up to physics (its anisotropic
    # constructor takes PHYSICS as an argument)
    assert [f[0] for f in orc._OParams._fields_] == [f[0] for f in dnhip._CParams._fields_][:-1]


def test_abi_version_and_error_text():
    lib = dnhip.load()
    assert lib.dn_abi_version() == 3
    assert isinstance(lib.dn_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    lib = dnhip.load()
    assert lib.dn_ctx_create(None, 0, None) == DN_ERR_ARG
    assert lib.dn_run_forward(None, 0) == DN_ERR_ARG
    assert lib.dn_get_seismogram(None, 0, None) == DN_ERR_ARG
    assert lib.dn_ctx_destroy(None) in (0, DN_ERR_ARG)
    assert lib.dn_last_error()

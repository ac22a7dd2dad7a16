"""CPU-side checks of the C ABI boundary: the library loads, exports every symbol
include/owt.h declares, and rejects bad case files before touching a device
(reference load_params, src/IO.cpp:8-74)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "owt.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(owt_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree(owt_mod):
    assert declared_symbols() == sorted(owt_mod.EXPORTS)


def test_library_exports_every_declared_symbol(owt_mod):
    L = ctypes.CDLL(owt_mod.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


REF_KEYS = """# physics
nu = 3.571428571428571e-4 # kinetic viscosity
dt = 1.0e-3
dkx = 0.5
dkz = 1.0
NX = 64
NY = 65
NZ = 64
alpha_mesh = 1.5
NT = 10
initial_condition = 1
PX = 1
PZ = 1
save_interval = 5
"""


def create_rc(owt_mod, text, tmp_path):
    p = tmp_path / "case.toml"
    p.write_text(text)
    L = owt_mod.lib()
    h = ctypes.c_void_p()
    opts = owt_mod._Opts(0, -1, -1, 0)
    rc = L.owt_create(os.fsencode(str(p)), ctypes.byref(opts), ctypes.byref(h))
    msg = L.owt_last_error(None).decode()
    if h:
        L.owt_destroy(h)
    return rc, msg


def test_missing_file_is_config_error(owt_mod):
    L = owt_mod.lib()
    h = ctypes.c_void_p()
    rc = L.owt_create(b"/nonexistent/case.toml", None, ctypes.byref(h))
    assert rc == 1 and not h
    assert "Error loading configuration file" in L.owt_last_error(None).decode()


@pytest.mark.parametrize("drop", ["nu", "dt", "dkx", "dkz", "NX", "NY", "NZ", "alpha_mesh", "NT", "PX", "PZ", "save_interval"])
def test_every_reference_key_is_required(owt_mod, tmp_path, drop):
    text = "\n".join(l for l in REF_KEYS.splitlines() if not l.startswith(drop + " "))
    rc, msg = create_rc(owt_mod, text, tmp_path)
    assert rc == 1 and drop in msg


@pytest.mark.parametrize("line,code", [("NX = 63", 2), ("NY = 5", 2), ("NX = 70", 2), ("dt = \"x\"", 1),
                                       ("NY = 1.5", 1), ("dt = 0.0", 2), ("NY = 300", 2)])
def test_bad_values_are_rejected(owt_mod, tmp_path, line, code):
    key = line.split("=")[0].strip()
    text = "\n".join(line if l.startswith(key + " ") else l for l in REF_KEYS.splitlines())
    rc, msg = create_rc(owt_mod, text, tmp_path)
    assert rc == code, msg


def test_reference_case_files_parse(owt_mod, tmp_path):
    # the key set of the reference's test-cases/*/case.toml (sizes reduced to what
    # the CPU check reaches: parsing succeeds and only device creation can fail)
    rc, msg = create_rc(owt_mod, REF_KEYS + "\n[owt]\nprecision = \"mixed\"\nmethod = \"plain\"\n", tmp_path)
    assert rc in (0, 3), msg
    if rc == 3:
        assert "HIP" in msg or "device" in msg


@pytest.mark.parametrize("extra,code", [("initial_condition = 4", 1), ("initial_condition = 0", 1),
                                        ("[owt]\ninit = \"bogus\"", 1), ("[owt]\nrestart = 5", 1),
                                        ("initial_condition = 3", None), ("[owt]\ninit = \"laminar\"", None)])
def test_initial_condition_keys(owt_mod, tmp_path, extra, code):
    """initial_condition (test-cases/*/case.toml:12) must be 1, 2 or 3; [owt] init one
    of the four names; valid files parse (only device creation can fail here)."""
    text = "\n".join(l for l in REF_KEYS.splitlines() if not l.startswith("initial_condition"))
    rc, msg = create_rc(owt_mod, text + "\n" + extra + "\n", tmp_path)
    if code is None:
        assert rc in (0, 3), msg
    else:
        assert rc == code, msg

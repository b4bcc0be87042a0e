"""CPU-side checks of the C ABI: the library loads, exports every entry point that
include/wqc_tn.h declares, and the host-only parts (backend seam) behave like the
reference's tn_backend_from_env (src/tn/backend.rs:12-29). No compute call is made
without a GPU."""

import ctypes as ct
import os
import re

import pytest

import wqc_tn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "wqc_tn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wqc_tn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = wqc_tn.load()
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(wqc_tn.EXPORTED) == decl


def test_backend_from_env(monkeypatch):
    lib = wqc_tn.load()
    for val, want in [(None, "cpu"), ("cpu", "cpu"), ("hip", "hip"), ("GPU", "hip"),
                      ("rocm", "hip"), ("webgpu", "cpu"), ("bogus", "cpu")]:
        if val is None:
            monkeypatch.delenv("WQC_TN_BACKEND", raising=False)
        else:
            monkeypatch.setenv("WQC_TN_BACKEND", val)
        assert lib.wqc_tn_backend_from_env().decode() == want


def test_null_arguments_are_rejected_without_a_device():
    lib = wqc_tn.load()
    out = ct.c_void_p()
    # null device handle: INVALID_ARG before any device call
    assert lib.wqc_tn_state_new(None, 4, 8, 0, ct.byref(out)) == wqc_tn.INVALID_ARG
    assert "null" in lib.wqc_tn_last_error().decode()
    assert lib.wqc_tn_apply_gate(None, None) == wqc_tn.INVALID_ARG
    assert lib.wqc_tn_backend_label(None).decode() == "hip"
    assert lib.wqc_tn_peak_vram_bytes(None) == 0


def test_gate_struct_layout_matches_header():
    assert ct.sizeof(wqc_tn.Gate) == 24
    assert ct.sizeof(wqc_tn.Stats) == 9 * 8 + 2 * 8 + 4 * 8

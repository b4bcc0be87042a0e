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
    assert ct.sizeof(wqc_tn.Stats) == 9 * 8 + 2 * 8 + 5 * 8


# ------------------------------------------------------------------ the Rust binding
C_TO_RUST = {"uint64_t": "u64", "double": "c_double", "int32_t": "i32", "size_t": "usize",
             "int": "c_int"}


def header_structs():
    text = open(os.path.join(ROOT, "include", "wqc_tn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct \w+ \{(.*?)\}\s*(\w+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            ctype, fname = decl.rsplit(None, 1)
            fields.append((fname, C_TO_RUST[ctype]))
        out[name] = fields
    return out


GUIDE = os.path.join(ROOT, "INTEGRATION.md")  # the integration guide with the Rust binding


def rust_block():
    text = open(GUIDE).read()
    m = re.search(r"## 2\. FFI declarations.*?```rust\n(.*?)```", text, flags=re.S)
    assert m, "the integration guide has no FFI declaration block"
    return m.group(1)


def rust_structs():
    out = {}
    for name, body in re.findall(r"pub struct (\w+)\s*\{(.*?)\}", rust_block(), flags=re.S):
        fields = re.findall(r"pub (\w+):\s*(\w+)", body)
        if fields:
            out[name] = fields
    return out


def test_rust_structs_match_header():
    """Every #[repr(C)] struct of the integration guide's Rust binding has the header's fields, in
    order and with the matching width: a missing field would let wqc_tn_get_stats write past a
    Rust caller's struct."""
    h, r = header_structs(), rust_structs()
    assert {"wqc_tn_stats", "wqc_tn_gate", "wqc_tn_slice"} <= set(h)
    for name, fields in h.items():
        assert r.get(name) == fields, (name, r.get(name), fields)
    assert ct.sizeof(wqc_tn.Stats) == 8 * len(h["wqc_tn_stats"])
    assert [f for f, _ in wqc_tn.Stats._fields_] == [f for f, _ in h["wqc_tn_stats"]]


def test_rust_functions_match_header():
    """Every entry point the header declares has a Rust declaration with the same parameter
    count."""
    text = open(os.path.join(ROOT, "include", "wqc_tn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    hdr = {}
    for name, params in re.findall(r"\b(wqc_tn_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        params = params.strip()
        hdr[name] = 0 if params in ("", "void") else params.count(",") + 1
    rs = {}
    for name, params in re.findall(r"pub fn (wqc_tn_[a-z0-9_]+)\s*\(([^)]*)\)", rust_block(), flags=re.S):
        params = params.strip().rstrip(",")
        rs[name] = 0 if not params else params.count(",") + 1
    assert hdr == rs, (set(hdr) ^ set(rs), {k: (hdr[k], rs.get(k)) for k in hdr if hdr[k] != rs.get(k)})

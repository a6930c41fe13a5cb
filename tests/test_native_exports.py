"""The C-ABI library loads and exports every entry point include/crux_lik.h
declares (no device calls)."""

import os
import re

from crux_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "crux_lik.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crux_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = nat.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(header_symbols()) == set(nat.SIGNATURES)


def test_error_strings():
    L = nat.lib()
    for code in range(8):
        assert L.crux_strerror(code)

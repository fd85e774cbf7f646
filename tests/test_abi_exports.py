"""The C-ABI library loads and exports every symbol include/*.h declares
(CPU only: no compute call is made without a GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

from numfast_amd import native

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    syms = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        syms += re.findall(r"\b(nfh?_[a-z0-9_]+)\s*\(", text)
    return sorted(set(syms))


@pytest.fixture(scope="module")
def so():
    path = Path(native.DEFAULT_LIB)
    if not path.exists():
        from numfast_amd import build
        build.build_lib()
    return ctypes.CDLL(str(path))


def test_header_declares_reference_surface():
    syms = declared_symbols()
    eager = [name for name, _, _ in native.EAGER]
    for name in eager:
        assert name in syms
    for g in ("variant", "mixed", "owner", "fused", "unique", "sort"):
        for name, _, _ in native.GROUPS[g]:
            assert name in syms, name


def test_library_exports_every_declared_symbol(so):
    missing = [s for s in declared_symbols() if not hasattr(so, s)]
    assert not missing, missing


def test_binding_binds_every_group(so):
    groups = native.bind(so)
    assert groups == set(native.GROUPS)


def test_no_gpu_means_device_error_not_abort(so):
    """Without a GPU the library loads and reports -6 instead of aborting."""
    native.bind(so)
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert so.nfh_device_ok() == 0
    import numpy as np
    k = np.arange(8, dtype=np.int32)
    p = np.zeros(8, np.int32)
    t = np.zeros(8, np.uint32)
    assert so.nf_sort_perm_i32(k.ctypes.data, 8, 0, p.ctypes.data, t.ctypes.data,
                               t.ctypes.data, p.ctypes.data) == -6
    # argument checks run before any device work
    assert so.nf_sort_perm_i32(k.ctypes.data, 8, 0, None, t.ctypes.data, t.ctypes.data,
                               p.ctypes.data) == -1
    assert so.nf_group_variant_i32(k.ctypes.data, k.ctypes.data, 8, 1, k.ctypes.data,
                                   p.ctypes.data, 8) == -3


def test_loader_env_contract(monkeypatch):
    """NUMFAST_NATIVE_DISABLE / NUMFAST_NATIVE_DLL behave like native_cpu.py:93-143,
    except that a missing library raises instead of falling back."""
    monkeypatch.setenv("NUMFAST_NATIVE_DISABLE", "1")
    assert not native.available() and native.why() == "disabled-by-env"
    with pytest.raises(native.NativeUnavailable):
        native.lib()
    monkeypatch.delenv("NUMFAST_NATIVE_DISABLE")
    monkeypatch.setenv("NUMFAST_NATIVE_DLL", "/nonexistent/libx.so")
    assert not native.available()
    assert native.why().startswith("unavailable:")
    monkeypatch.setenv("NUMFAST_NATIVE_DLL", native.DEFAULT_LIB)
    assert native.available()

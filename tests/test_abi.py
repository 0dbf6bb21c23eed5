"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its host-side input generators agree with the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "yafel_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(yafel_hip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from yafel_amd import hip
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(hip.lib, s)]
    assert not missing, missing


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    import importlib.util
    src = os.path.join(ROOT, "yafel_amd", "hip.py")
    fake = tmp_path / "yafel_fake"
    fake.mkdir()
    (fake / "hip.py").write_text(open(src).read())
    spec = importlib.util.spec_from_file_location("yafel_fake_hip", str(fake / "hip.py"))
    mod = importlib.util.module_from_spec(spec)
    with pytest.raises(ImportError):
        spec.loader.exec_module(mod)


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_structured_mesh_matches_oracle(orc, dim, p):
    from yafel_amd import hip
    nodes, cells = hip.structured_mesh(dim, 5, 4, 3 if dim == 3 else 0, p)
    on, oc = orc.structured_mesh(dim, 5, 4, 3 if dim == 3 else 0, 0 if p == 1 else 1)
    assert np.array_equal(nodes, on) and np.array_equal(cells, oc)


def test_normal_fill_matches_oracle(orc):
    # same splitmix64 stream and Box-Muller map; the host libm's sin/cos/log may
    # differ in the last ulp between compilers, so parity tests share one input array
    from yafel_amd import hip
    a, b = hip.normal(123, 100001, 0.5, 2.0), orc.normal(123, 100001, 0.5, 2.0)
    assert np.abs(a - b).max() <= 4e-16 * np.abs(b).max()


def test_error_reporting_without_device():
    from yafel_amd import hip
    with pytest.raises(hip.YafelHipError):
        hip.structured_mesh(4, 1, 1, 1, 1)
    assert b"dim" in hip.lib.yafel_hip_last_error(None)

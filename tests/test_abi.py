"""C ABI and host-side logic, without a GPU: the library loads, exports every declared symbol,
reports the missing device, validates arguments, and the storage lowering mirrors the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

import particular_amd as pa
from particular_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(pnb_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def _c_params(decl):
    """Parameter count of each C declaration `int pnb_x(...)` in the header."""
    txt = re.sub(r"/\*.*?\*/", "", decl, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(pnb_[a-z0-9_]+)\s*\(([^)]*)\)", txt):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_rust_bindings_match_header():
    """Every function the Rust crate declares (particular-hip/src/ffi.rs) is declared in the
    header with the same number of parameters and exported by the library."""
    header = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            header.update(_c_params(open(os.path.join(ROOT, "include", fn)).read()))
    ffi = open(os.path.join(ROOT, "particular-hip", "src", "ffi.rs")).read()
    rust = {name: (0 if not params.strip() else params.count(",") + 1 - params.rstrip().endswith(","))
            for name, params in re.findall(r"pub fn (pnb_[a-z0-9_]+)\(([^)]*)\)", ffi)}
    assert len(rust) >= 10
    L = _lib.lib()
    for name, count in rust.items():
        assert name in header, name
        assert header[name] == count, (name, header[name], count)
        assert hasattr(L, name), name


def gpu_present():
    return pa.device_count() > 0


@pytest.mark.skipif(gpu_present(), reason="checks the no-device path")
def test_no_device_is_reported():
    L = _lib.lib()
    n = ctypes.c_int(5)
    assert L.pnb_device_count(ctypes.byref(n)) == _lib.PNB_E_NODEV and n.value == 0
    h = ctypes.c_void_p()
    assert L.pnb_create(0, ctypes.byref(h)) == _lib.PNB_E_NODEV and not h.value
    with pytest.raises(pa.PnbError) as e:
        pa.HipResources(0)
    assert e.value.status == _lib.PNB_E_NODEV


def test_null_context_is_an_argument_error():
    L = _lib.lib()
    assert L.pnb_bf_accel_f32(None, None, 0, None, 0, 0.0, 1, None) == _lib.PNB_E_ARG
    assert L.pnb_bf_accel_f32_dev(None, None, 0, None, 0, 0.0, 1, None) == _lib.PNB_E_ARG
    assert L.pnb_reordered_bf_accel_f32(None, None, 0, 0.0, 1, None) == _lib.PNB_E_ARG
    assert L.pnb_synchronize(None) == _lib.PNB_E_ARG
    assert L.pnb_set_stream(None, None) == _lib.PNB_E_ARG
    assert L.pnb_reset_stream(None) == _lib.PNB_E_ARG
    assert L.pnb_profile_enable(None, 1) == _lib.PNB_E_ARG
    assert L.pnb_direct_f64_dev(None, None, 0, None, 0, 0.0, None) == _lib.PNB_E_ARG
    assert L.pnb_create(0, None) == _lib.PNB_E_ARG
    assert L.pnb_device_count(None) == _lib.PNB_E_ARG
    L.pnb_destroy(None)  # no-op


def test_status_strings():
    L = _lib.lib()
    for s in range(0, 7):
        assert L.pnb_status_string(s).decode()
    assert L.pnb_status_string(99).decode() == "unknown status"


def test_ordered_is_a_stable_partition():
    p = np.array([[0, 0, 0, 0], [1, 0, 0, 2], [2, 0, 0, 0], [3, 0, 0, 5], [4, 0, 0, 1]], np.float32)
    o = pa.Ordered.new(p)
    assert o.affecting_len() == 3
    np.testing.assert_array_equal(o.affecting()[:, 0], [1, 3, 4])
    np.testing.assert_array_equal(o.non_affecting()[:, 0], [0, 2])
    r = pa.Reordered.new(p)
    assert r.unordered is p
    np.testing.assert_array_equal(r.affecting(), o.affecting())
    np.testing.assert_array_equal(r.reordered(), o.particles())


def test_storage_lowering_mirrors_reference():
    p = np.zeros((3, 4), np.float32)
    q = np.ones((2, 4), np.float32)
    assert pa._lower(p)[0] == "between" and pa._lower(p)[1] is p and pa._lower(p)[2] is p
    k, a, b = pa._lower(pa.Between(p, q))
    assert a is p and b is q
    o = pa.Ordered.new(q)
    k, a, b = pa._lower(o)
    assert a is o.particles()
    assert pa._lower(pa.Reordered.new(p))[0] == "reordered"
    k, a, b = pa._lower(pa.Reordered.new(q, lambda x: x[0] > 0))
    assert k == "between" and a is q and b.shape == (2, 4)


def test_interaction_constructors():
    assert pa.Acceleration.checked() == pa.Acceleration(True, 0.0)
    assert not pa.Acceleration.unchecked().is_checked
    s = pa.AccelerationSoftened.checked(1e-3)
    assert s.softening == pytest.approx(1e-3) and s.is_checked
    assert not pa.AccelerationSoftened.unchecked(0.5).is_checked


def test_mlp_parameter_layout_matches_library():
    from particular_amd import learned

    L = _lib.lib()
    assert L.pnb_mlp_num_params() == learned.IN * 64 + 64 + 64 * 64 + 64 + 64 * learned.OUT + learned.OUT
    flat = np.random.default_rng(0).normal(size=learned.num_params()).astype(np.float32)
    back = learned.dict_to_params(learned.params_to_dict(flat))
    np.testing.assert_array_equal(back, flat)
    assert not learned.quadrupole_params().any()  # zero output layer: the residual alone


def test_tree_node_record_size():
    # pnb_tree_node: com[3], mass, mass64 (8-aligned), width, skip, begin, count, level, internal,
    # quad[6] -> 72 bytes
    assert _lib.tree_node_dtype().itemsize == 72

"""The C-ABI library loads without a GPU and exports every entry point that
include/nbg.h declares; without a usable device the product path raises
instead of falling back to the CPU."""
import ctypes
import os
import re

import pytest

import nbodygradient_amd as nb
from nbodygradient_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(nbg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("nbg_ctx_create", "nbg_ahl21_step", "nbg_integrate_steps", "nbg_integrate_to",
                     "nbg_integrate_transits", "nbg_batch_create", "nbg_batch_run", "nbg_batch_download"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnbg_hip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, "declared in include/nbg.h but not exported: %s" % missing


def test_no_device_raises():
    """On a host without a usable GPU the context cannot be created: the
    integration entry points have no CPU fallback."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnbg_hip.so not built")
    try:
        ctx = nb.Context(0)
    except nb.NbgError as e:
        assert "no usable HIP device" in str(e)
        return
    ctx.close()
    pytest.skip("a GPU is present; the no-device path is not reachable here")


def test_server_footprint_matches_the_record_layout():
    """Device bytes of one served request: state, Jacobians (J, Jerr,
    jac_init), the transit-step Jacobian copy (and, from n = 7 on, the working
    J/Jerr in HBM), tt, dtdq0, dtdelements, counters and the pair mask, rounded
    up to 256 bytes (no GPU needed)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnbg_hip.so not built")
    L = _lib.lib()
    for n, ntbv, ntt in [(2, 1, 5), (3, 1, 103), (9, 1, 2003), (10, 3, 40)]:
        ws = (168 if n >= 7 else 84) * n * n
        b = 8 * ((21 * n + 4) + 3 * 49 * n * n + ws + ntbv * n * ntt + 2 * ntbv * n * ntt * 7 * n + (n + 3))
        b += (n * n + 7) // 8 * 8
        assert L.nbg_server_footprint(n, ntbv, ntt) == (b + 255) // 256 * 256


def test_autotune_candidates_start_with_the_default():
    """The autotuner's candidate table (host code, no GPU needed): its first
    entry is the context's default (automatic two-wave selection of at most
    128 systems, workgroups per CU as LDS allows), and it also tries one wave
    per system throughout."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnbg_hip.so not built")
    L = _lib.lib()
    nc = L.nbg_batch_autotune_candidate(-1, None, None, None)
    assert nc >= 2
    cands = []
    for k in range(nc):
        a, b, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        assert L.nbg_batch_autotune_candidate(k, ctypes.byref(a), ctypes.byref(b), ctypes.byref(g)) == nc
        cands.append((a.value, b.value, g.value, L.nbg_batch_autotune_systems_per_wave(k)))
    assert cands[0] == (1, 128, 0, 1)
    assert any(c[0] == 0 for c in cands)
    assert {c[3] for c in cands} == {1, 2}  # one and two systems per wave are both tried
    assert L.nbg_batch_autotune_systems_per_wave(nc) == -1
    assert len(set(cands)) == nc

"""C-ABI checks that need no GPU: the library loads, exports every function include/gatk_pairhmm.h
declares, and reports the missing device the way VectorLoglessPairHMM expects (load() false ->
HardwareFeatureException, VectorLoglessPairHMM.java:63-66)."""
import ctypes
import os
import re

import pytest

from gatk_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gatk_pairhmm.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pairhmm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists():
    assert _declared_functions() == sorted(_native.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name


def test_status_codes_match_header():
    text = open(HEADER).read()
    codes = dict((k, int(v)) for k, v in re.findall(r"(PAIRHMM_(?:OK|ERR_[A-Z_]+))\s*=\s*(\d+)", text))
    assert codes == {"PAIRHMM_OK": _native.PAIRHMM_OK, "PAIRHMM_ERR_NO_DEVICE": _native.PAIRHMM_ERR_NO_DEVICE,
                     "PAIRHMM_ERR_BAD_ARG": _native.PAIRHMM_ERR_BAD_ARG, "PAIRHMM_ERR_OOM": _native.PAIRHMM_ERR_OOM,
                     "PAIRHMM_ERR_KERNEL": _native.PAIRHMM_ERR_KERNEL, "PAIRHMM_ERR_STATE": _native.PAIRHMM_ERR_STATE}


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_raises_hardware_feature_exception():
    from gatk_amd.pairhmm import HardwareFeatureException, HipLoglessPairHMM, HipPairHmm

    assert HipPairHmm().load(None) is False
    with pytest.raises(HardwareFeatureException):
        HipLoglessPairHMM()
    lib = _native.load_library()
    ctx = ctypes.c_void_p()
    args = _native.NativeArguments(1, 0)
    assert lib.pairhmm_initialize(ctypes.byref(args), ctypes.byref(ctx)) == _native.PAIRHMM_ERR_NO_DEVICE
    assert not ctx
    assert "gfx950" in _native.last_error()


def test_null_context_is_a_state_error():
    lib = _native.load_library()
    assert lib.pairhmm_compute_likelihoods(None, None, 1, None, 1, None) == _native.PAIRHMM_ERR_STATE
    assert lib.pairhmm_done(None) == _native.PAIRHMM_OK
    assert lib.pairhmm_batch_free(None) == _native.PAIRHMM_OK

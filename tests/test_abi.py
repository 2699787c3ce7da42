"""The C-ABI library loads and exports every symbol the public headers declare (no GPU calls)."""

import ctypes
import re
from pathlib import Path

import pytest

from rcabeam_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols() -> set[str]:
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names.update(re.findall(r"\b(rcabeam_[a-z0-9_]+)\s*\(", text))
    return names


def test_headers_declare_reference_symbols():
    """The 15 entry points of the reference header (include/rcabeam.h:15-235) are all declared."""
    reference = {
        "rcabeam_status_string",
        "rcabeam_opw_device",
        "rcabeam_xdoppler_pd_device",
        "rcabeam_rc_fmas_pd_device",
        "rcabeam_delay_rca_channels_device",
        "rcabeam_delay_rca_channel_data_device",
        "rcabeam_opw_pd_from_channels_device",
        "rcabeam_xdoppler_pd_from_channels_device",
        "rcabeam_rc_fmas_pd_from_channels_device",
        "rcabeam_fast_pd_from_channels_device",
        "rcabeam_dmas_ccf_acf_from_channels_device",
        "rcabeam_pack_ensemble_device",
        "rcabeam_ensemble_device",
        "rcabeam_pack_ensemble_fp16_device",
        "rcabeam_opw_fp16_device",
    }
    assert reference <= _declared_symbols()


@pytest.mark.skipif(not _lib.LIBRARY_PATH.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIBRARY_PATH))
    missing = [name for name in sorted(_declared_symbols()) if not hasattr(lib, name)]
    assert not missing
    # Every symbol bound in Python is declared in a header.
    assert set(_lib.EXPORTED) <= _declared_symbols()


@pytest.mark.skipif(not _lib.LIBRARY_PATH.exists(), reason="library not built")
def test_status_strings_and_argument_errors_without_gpu_work():
    lib = _lib.load(require_gpu=False)
    assert _lib.status_string(0) == "success"
    assert _lib.status_string(2) == "invalid argument"
    # Null pointers are rejected before any device call.
    assert lib.rcabeam_pack_ensemble_device(None, None, 4, 2, 2, 2) == _lib.ERROR_ARGUMENT
    assert lib.rcabeam_opw_device(None, None, 1, 1, 1) == _lib.ERROR_ARGUMENT
    # Subspace iteration: k in [1, 12], k < frames, k x frames <= 4096, 1 <= iterations <= 1000.
    fake = ctypes.c_void_p(4096)  # never dereferenced: the shape checks come first
    assert lib.rcabeam_dominant_subspace_device(None, 200, 4, 12, fake) == _lib.ERROR_ARGUMENT
    for frames, k, iterations in ((200, 0, 12), (200, 13, 12), (4, 4, 12), (512, 9, 12), (200, 4, 0), (200, 4, 1001)):
        assert lib.rcabeam_dominant_subspace_device(fake, frames, k, iterations, fake) == _lib.ERROR_ARGUMENT
    assert lib.rcabeam_dominant_subspace_device(ctypes.c_void_p(4104), 200, 4, 12, fake) == _lib.ERROR_ARGUMENT
    with pytest.raises(ValueError, match="invalid argument"):
        _lib.check(_lib.ERROR_ARGUMENT, "probe")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.ERROR_DEVICE, "probe")

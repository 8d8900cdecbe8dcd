"""The C ABI: the library loads and exports every function include/esm.h declares (no GPU needed)."""
import ctypes as C
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int32_t|const char \*|void)\s*\*?\s*(esm_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "esm_create" in names and "esm_energy" in names and len(names) >= 8


def test_library_exports_every_declared_symbol():
    from esmsampling import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound, "every declared entry point has a ctypes signature"


def test_version_and_error_without_gpu_work():
    from esmsampling import _lib
    lib = _lib.load()
    assert lib.esm_version() == 9
    # an invalid argument is rejected on the host, before any device call
    h = C.c_void_p()
    assert lib.esm_create(None, 1, 64, C.byref(h)) == -1
    assert b"null" in lib.esm_last_error()


def test_batch_and_autotune_reject_null_arguments_without_gpu_work():
    from esmsampling import _lib
    lib = _lib.load()
    out = C.c_void_p()
    assert lib.esm_batch_create(None, 1, C.byref(out)) == -1 and b"null" in lib.esm_last_error()
    assert lib.esm_batch_sweeps(None, 1, None, None, 40.0, None) == -1
    assert lib.esm_batch_counts(None, None) == -1
    assert lib.esm_batch_destroy(None) == 0
    assert lib.esm_sweep_block(None, 128) == -1
    assert lib.esm_sweep_pad(None, 1) == -1
    assert lib.esm_autotune(None, 1, None, 0, 40.0, None, None, None) == -1
    assert lib.esm_score_measure(None, 1) == -1 and b"null" in lib.esm_last_error()
    assert lib.esm_heuristic_state(None) == -1 and b"null" in lib.esm_last_error()
    assert _lib.BATCH_MAX == 64

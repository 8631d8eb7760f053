"""The C-ABI library loads and exports every function include/farcur.h declares
(no GPU calls)."""
import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "farcur.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(?:int|void|double|const char\*)\s+\**\s*([a-z_][a-z0-9_]*)\s*\(", text))
    return sorted(names)


def test_header_declares_drop_in_and_native_api():
    names = _declared()
    assert "cuda_complete_compression" in names
    for n in ("farcur_create", "farcur_compress_block", "farcur_batch_compress", "farcur_fill", "farcur_pinv"):
        assert n in names


def test_library_exports_every_declared_symbol(farcur):
    lib = ctypes.CDLL(str(farcur.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_drop_in_arity():
    # cuda_complete_compression takes the reference's 66 arguments (CU:1564-1629).
    text = (ROOT / "include" / "farcur.h").read_text()
    m = re.search(r"double cuda_complete_compression\((.*?)\);", text, flags=re.S)
    assert m and len(m.group(1).split(",")) == 66


def test_host_only_calls_work_without_gpu(farcur):
    v = farcur.vtest(1, 2, 5)
    assert v.shape == (5,)
    assert farcur.lib().farcur_version().decode().startswith("farcur")

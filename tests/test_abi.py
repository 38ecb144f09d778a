"""The C-ABI library loads on a host without a GPU and exports every entry
point the public headers declare (no compute calls here)."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "nd-image-codecs_amd" / "libndic_hip.so"


def declared():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(ndic_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_boundary():
    names = declared()
    for required in ["ndic_htj2k_encode_chunk", "ndic_htj2k_decode_chunk", "ndic_htj2k_decode_plane",
                     "ndic_lift_forward_chunk", "ndic_lift_inverse_chunk", "ndic_lift_ht_decode_batch",
                     "ndic_lift_ht_encode_batch", "ndic_buffer_free", "ndic_version"]:
        assert required in names, required


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(LIB))
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string_without_gpu():
    lib = C.CDLL(str(LIB))
    lib.ndic_version.restype = C.c_char_p
    assert lib.ndic_version().decode().startswith("nd-image-codecs-amd")

#!/usr/bin/env python3
"""Collect the reference's committed golden data into ``tests/golden/``.

Everything written here is data the reference's own tests pin (inputs and
expected outputs), so the parity tests run without the reference tree:

- ``nd_lift_vectors.json``: ``fixtures/nd-lift/vectors.json`` (13 forward
  cases, read by ``crates/ndic-lift/tests/vectors.rs:20``);
- ``tiny-gradient-8x8.j2c``, ``tiny-gradient-8x8.jph``,
  ``tiny-chunk-4x8x8.ndht``: ``fixtures/codestreams/`` (byte-stability pins in
  ``crates/ndic-codestream/tests/fixtures.rs:25-73`` and
  ``crates/ndic-zarr/tests/htj2k_zarrs.rs:373-398``);
- ``zfp_checksums.json`` and ``tiny-chunk-4x8x8-rate8.zfp``: ``fixtures/zfp/``
  (``crates/ndic-zfp/tests/checksums.rs:21-86``);
- ``ht_tables.json``: the HT lookup tables the reference ships pre-generated
  (``crates/ndic-htj2k/src/block/tables_data.rs``: ``DEC_VLC_TBL0/1``,
  ``ENC_VLC_TBL0/1``, ``UVLC_TBL0/1``), as expected outputs for this project's
  own derivation of those tables from the T.814 rows.

Usage: make_golden.py [reference-root]
"""

from __future__ import annotations

import json
import re
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def u16_table(text: str, name: str) -> list[int]:
    m = re.search(r"pub static %s: \[u16; (\d+)\] = \[(.*?)\];" % name, text, re.S)
    if m is None:
        raise SystemExit(f"table {name} not found")
    vals = [int(v.strip(), 0) for v in m.group(2).split(",") if v.strip()]
    if len(vals) != int(m.group(1)):
        raise SystemExit(f"table {name}: expected {m.group(1)} entries, got {len(vals)}")
    return vals


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    fx = root / "fixtures"
    shutil.copyfile(fx / "nd-lift/vectors.json", HERE / "nd_lift_vectors.json")
    for name in ("tiny-gradient-8x8.j2c", "tiny-gradient-8x8.jph", "tiny-chunk-4x8x8.ndht"):
        shutil.copyfile(fx / "codestreams" / name, HERE / name)
    shutil.copyfile(fx / "zfp/checksums.json", HERE / "zfp_checksums.json")
    shutil.copyfile(fx / "zfp/tiny-chunk-4x8x8-rate8.zfp", HERE / "tiny-chunk-4x8x8-rate8.zfp")

    text = (root / "crates/ndic-htj2k/src/block/tables_data.rs").read_text()
    tables = {n: u16_table(text, n) for n in (
        "DEC_VLC_TBL0", "DEC_VLC_TBL1", "ENC_VLC_TBL0", "ENC_VLC_TBL1", "UVLC_TBL0", "UVLC_TBL1")}
    (HERE / "ht_tables.json").write_text(json.dumps(tables) + "\n")


if __name__ == "__main__":
    main()

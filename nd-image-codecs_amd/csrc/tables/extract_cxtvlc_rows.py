#!/usr/bin/env python3
"""Write ``ht_cxtvlc_rows.inc``: the ITU-T T.814 Annex C CxtVLC code tables as data.

The HT cleanup pass codes each quad's significance pattern with a context
dependent variable-length code whose codeword assignments are fixed by the
standard (T.814 Annex C, one table for the initial quad row and one for the
other rows).  Each row is ``{c_q, rho, u_off, e_k, e_1, cwd, cwd_len}``.

The rows are transcribed from the reference's generated table module
(``crates/ndic-htj2k/src/block/tables_data.rs``, statics ``RAW_TBL0`` and
``RAW_TBL1``), read as text.  Every lookup table the coder uses (decoder and
encoder CxtVLC, UVLC) is derived from these rows by this project's own code.

Usage: extract_cxtvlc_rows.py [reference-root] [output.inc]
"""

from __future__ import annotations

import re
import sys
from pathlib import Path


def parse_rows(text: str, name: str) -> list[tuple[int, ...]]:
    m = re.search(r"pub static %s: \[VlcSrcEntry; (\d+)\] = \[(.*?)\];" % name, text, re.S)
    if m is None:
        raise SystemExit(f"table {name} not found")
    rows = [tuple(int(x.strip(), 0) for x in e.split(","))
            for e in re.findall(r"e\(([^)]*)\)", m.group(2))]
    if len(rows) != int(m.group(1)) or any(len(r) != 7 for r in rows):
        raise SystemExit(f"table {name}: malformed rows")
    return rows


def emit(name: str, rows: list[tuple[int, ...]]) -> str:
    lines = [f"static const unsigned char {name}[{len(rows)}][7] = {{"]
    for r in rows:
        lines.append("  {" + ", ".join(str(v) for v in r) + "},")
    lines.append("};")
    return "\n".join(lines)


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = Path(sys.argv[2] if len(sys.argv) > 2 else Path(__file__).with_name("ht_cxtvlc_rows.inc"))
    text = (root / "crates/ndic-htj2k/src/block/tables_data.rs").read_text()
    body = [
        "// ITU-T T.814 Annex C CxtVLC code tables, one row per codeword:",
        "// {c_q, rho, u_off, e_k, e_1, cwd, cwd_len}.  Data only; written by",
        "// extract_cxtvlc_rows.py.",
        emit("kCxtVlcRows0", parse_rows(text, "RAW_TBL0")),
        emit("kCxtVlcRows1", parse_rows(text, "RAW_TBL1")),
        "",
    ]
    out.write_text("\n".join(body))


if __name__ == "__main__":
    main()

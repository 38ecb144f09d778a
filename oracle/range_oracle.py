"""Test infrastructure: a plain-Python restatement of the reference's byte-range
planner, the checker for ``ndic_range_plan_chunk`` and for the plane choice
of ``ndic_htj2k_decode_select_batch``.  Only tests/ import it.

Follows:
  * ``ChunkHeader::parse`` -- crates/ndic-codestream/src/container.rs:199-340
  * ``RangeIndex`` (resolution_for, header_range, plane_prefix_range,
    thumbnail, plane, thumbnail_3d, coalesce) --
    crates/ndic-codestream/src/range.rs:100-369
  * ``approx_len`` / ``low_pass_indices`` -- crates/ndic-lift/src/lib.rs:195-225
    (band count: crates/ndic-lift/src/kernel.rs, levels applied while the
    running length is at least 2)
  * the zarrs partial decoder's plane walk --
    crates/ndic-zarr/src/htj2k_codec.rs:376-387

Pinned by tests/test_range_plans.py: the reference's low_pass_indices
vectors (lib.rs:297-315), the container sample of container.rs:349-364, and
hand-derived plans for it.
"""

from __future__ import annotations

import struct

KINDS = {"delta": 0, "haar": 1, "lift53": 2}


class PlanError(Exception):
    def __init__(self, kind: str, message: str, offset: int = 0):
        super().__init__(message)
        self.kind, self.message, self.offset = kind, message, offset


def parse_chunk_prefix(b: bytes) -> dict:
    def err(off, msg):
        raise PlanError("Codestream", msg, off)
    if len(b) < 8:
        err(0, "chunk header prefix too short")
    if b[:4] != b"ndht":
        err(0, "not an htj2k chunk (bad magic)")
    if b[4] != 1:
        err(4, "unsupported htj2k chunk container version")
    flags, lv, ndim = b[5], b[6], b[7]
    if flags & ~1:
        err(5, "unknown htj2k chunk container flags")
    if lv > 32:
        err(6, "xy_levels exceeds the 32-level J2K bound (T.800 SPcod)")
    if not 2 <= ndim <= 32:
        err(7, "chunk ndim outside 2..=32")
    fixed = 8 + 4 * ndim
    if len(b) < fixed:
        err(len(b), "truncated chunk header")
    dims = []
    for i in range(ndim):
        d = struct.unpack_from("<I", b, 8 + 4 * i)[0]
        if d == 0:
            err(8 + 4 * i, "zero chunk dimension")
        dims.append(d)
    np_ = 1
    for d in dims[:-2]:
        np_ *= d
        if np_ > 1 << 24:
            err(8, "chunk declares more planes than supported")
    out = {"dims": dims, "xy_levels": lv, "header_len": fixed, "planes": None}
    if not flags & 1:
        return out
    entry = 12 + 4 * (lv + 1)
    need = fixed + np_ * entry
    if len(b) < need:
        err(len(b), "truncated coefficient-plane index")
    planes, min_off = [], need
    for i in range(np_):
        at = fixed + i * entry
        off, ln = struct.unpack_from("<QI", b, at)
        prefix = list(struct.unpack_from(f"<{lv + 1}I", b, at + 12))
        if off < min_off:
            err(at, "plane offsets overlap or precede the index")
        if any(prefix[k] > prefix[k + 1] for k in range(lv)) or prefix[-1] > ln:
            err(at, "plane prefix lengths are inconsistent")
        min_off = off + ln
        planes.append((off, ln, prefix))
    out["planes"], out["header_len"] = planes, need
    return out


def level_dims(w: int, h: int, levels: int):
    for _ in range(levels):
        w, h = (w + 1) // 2, (h + 1) // 2
    return w, h


def resolution_for(w: int, h: int, levels: int, max_px: int):
    best = (0, level_dims(w, h, levels))
    for r in range(levels + 1):
        dims = level_dims(w, h, levels - r)
        if max(dims) <= max_px:
            best = (r, dims)
    return best[0], best[1][0], best[1][1]


def band_count(n: int, levels: int) -> int:
    c, m = 0, n
    while c < levels and m >= 2:
        c, m = c + 1, (m + 1) // 2
    return c


def approx_len(kind: int, group_len: int, levels: int) -> int:
    if kind == 0:
        return min(group_len, 1)
    a = band_count(group_len, levels)
    return -(-group_len // (1 << a))


def low_pass_indices(step: dict, extent: int) -> list:
    if extent == 0:
        return []
    group = step.get("group", 0) or extent
    group = min(group, extent)
    out, base = [], 0
    kind = KINDS[step["kind"]]
    while base < extent:
        n = min(group, extent - base)
        out.extend(range(base, base + approx_len(kind, n, step.get("levels", 1))))
        base += n
    return out


def coalesce(ranges):
    out = []
    for s, e in sorted(ranges):
        if out and s <= out[-1][1] + 1:
            out[-1] = (out[-1][0], max(out[-1][1], e))
        else:
            out.append((s, e))
    return out


def _prefix_range(hdr, plane, max_res):
    planes = hdr["planes"]
    if plane >= len(planes):
        raise PlanError("InvalidArgument", f"plane {plane} out of range for a {len(planes)}-plane chunk")
    off, _, prefix = planes[plane]
    n = prefix[min(max_res, hdr["xy_levels"])]
    if n == 0:
        raise PlanError("InvalidArgument", "plane declares an empty resolution prefix")
    return (off, off + n - 1)


def thumbnail3d_planes(dims, steps):
    lead = dims[:-2]
    for st in steps:
        if st["dimension"] >= len(lead):
            raise PlanError("InvalidArgument",
                            f"nd_lift step on dimension {st['dimension']} is not a leading chunk axis "
                            f"(chunk has {len(lead)} leading dimensions)")
    masks = []
    for d, extent in enumerate(lead):
        st = next((s for s in steps if s["dimension"] == d), None)
        if st is None:
            masks.append([True] * extent)
        else:
            m = [False] * extent
            for i in low_pass_indices(st, extent):
                m[i] = True
            masks.append(m)
    n = 1
    for e in lead:
        n *= e
    sel = []
    for lin in range(n):
        rem, keep = lin, True
        for m in reversed(masks):
            keep &= m[rem % len(m)]
            rem //= len(m)
        if keep:
            sel.append(lin)
    return sel, [sum(m) for m in masks]


def plan(prefix: bytes, target: str, arg: int, steps=()) -> dict:
    hdr = parse_chunk_prefix(prefix)
    if hdr["planes"] is None:
        raise PlanError("Unsupported", "chunk was written without a coefficient-plane index (index: false)")
    dims = hdr["dims"]
    W, H, L = dims[-1], dims[-2], hdr["xy_levels"]
    ranges = [(0, hdr["header_len"] - 1)] if hdr["header_len"] else []
    if target in ("thumbnail", "thumbnail_3d"):
        max_res, w, h = resolution_for(W, H, L, arg)
        if target == "thumbnail":
            ranges.append(_prefix_range(hdr, 0, max_res))
            planes, size = [0], [h, w]
        else:
            planes, counts = thumbnail3d_planes(dims, list(steps))
            ranges += [_prefix_range(hdr, p, max_res) for p in planes]
            size = counts + [h, w]
    else:
        if arg >= len(hdr["planes"]):
            raise PlanError("InvalidArgument", f"plane {arg} out of range for a {len(hdr['planes'])}-plane chunk")
        off, ln, _ = hdr["planes"][arg]
        if ln:
            ranges.append((off, off + ln - 1))
        planes, size, max_res = [arg], [H, W], L
    ranges = coalesce(ranges)
    return {"target": target, "decoded_size": size, "max_res": max_res, "planes": planes, "ranges": ranges,
            "total_bytes": sum(e - s + 1 for s, e in ranges)}


def subset_planes(shape, start, count) -> list:
    """The planes an axis-aligned subset touches, C order of the leading dims
    (htj2k_codec.rs:376-387)."""
    planes = [0]
    lead = list(shape[:-2])
    for d in range(len(lead)):
        extent = 1
        for e in lead[d + 1:]:
            extent *= e
        planes = [b + (start[d] + i) * extent for b in planes for i in range(count[d])]
    return planes


def plane_codestream(chunk: bytes, plane: int) -> bytes:
    """Plane `plane`'s codestream bytes, located through the index."""
    hdr = parse_chunk_prefix(chunk)
    off, ln, _ = hdr["planes"][plane]
    return chunk[off:off + ln]

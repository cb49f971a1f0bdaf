"""The scan_csv_v1 format of the reference's scripts/utils/scan_csv.jl
(ScanCSV: write_metadata, write_header, read_existing_keys):

* a line starting with `#` is metadata / comment (`# key: value`);
* the first other line is the CSV header;
* keys are tuples of Float64 read from named columns, for resuming a scan.

Plus `julia_string`, which formats a Float64 as Julia's `string(x)` does
(shortest round-trip digits; `1.0e-5`, `NaN`, `Inf`), so rows written here
read like the reference's.
"""
from __future__ import annotations

import math
import os


def write_metadata(f, metadata):
    """ScanCSV.write_metadata  scan_csv.jl:17-22: `# key: value` per item."""
    for k, v in metadata.items():
        f.write(f"# {k}: {v}\n")


def write_header(f, cols):
    """ScanCSV.write_header  scan_csv.jl:25-27."""
    f.write(",".join(cols) + "\n")


def _is_comment(line):
    s = line.strip()
    return not s or s.startswith("#")


def read_existing_keys(path, key_cols):
    """ScanCSV.read_existing_keys  scan_csv.jl:52-92: the set of key tuples
    (floats of `key_cols`) of the data rows; an unknown column, a short row
    or an unparsable value skips as the reference does."""
    keys = set()
    if not os.path.isfile(path):
        return keys
    with open(path) as f:
        header = None
        for line in f:
            if _is_comment(line):
                continue
            header = line
            break
        if header is None:
            return keys
        colmap = {c.strip(): i for i, c in enumerate(header.strip().split(","))}
        if any(c not in colmap for c in key_cols):
            return keys
        idxs = [colmap[c] for c in key_cols]
        for line in f:
            if _is_comment(line):
                continue
            parts = line.split(",")
            if len(parts) < max(idxs) + 1:
                continue
            try:
                keys.add(tuple(float(parts[i].strip()) for i in idxs))
            except ValueError:
                continue
    return keys


def julia_string(x):
    """Julia's string(::Float64): shortest digits that round-trip, a `.0` on
    integral mantissas, exponents without padding (`1.0e-5`, `2.5e20`)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    s = repr(x)
    if "e" not in s:
        if abs(x) < 1e6:
            return s
        # Julia writes magnitudes from 1e6 up in exponent form
        sign = "-" if s.startswith("-") else ""
        ip, fp = s.lstrip("-").split(".")
        digits = (ip + fp).rstrip("0") or "0"
        mant = digits[0] + "." + (digits[1:] or "0")
        return f"{sign}{mant}e{len(ip) - 1}"
    mant, exp = s.split("e")
    if "." not in mant:
        mant += ".0"
    return f"{mant}e{int(exp)}"


def julia_bool(b):
    return "true" if b else "false"


__all__ = ["write_metadata", "write_header", "read_existing_keys", "julia_string", "julia_bool"]

"""Embed the CIE 1931 CMF and D65 / A SPD CSVs into the library as C string
literals (the reference embeds the same files with include_str!,
spectral-core/src/cie.rs:6-8).  Usage: gen_tables.py <data_dir> <out.inc>"""
import sys
from pathlib import Path


def main() -> None:
    data, out = Path(sys.argv[1]), Path(sys.argv[2])
    parts = []
    for name, sym in (("cie_1931_2deg.csv", "kCmfCsv"), ("illuminant_d65.csv", "kD65Csv"),
                      ("illuminant_a.csv", "kACsv")):
        text = (data / name).read_text()
        # the device's table lookup (cmf_at, ill_at) blends the last row with
        # itself at fraction 0, which returns every value but -0 unchanged
        for line in text.splitlines():
            for field in line.split(",")[1:]:
                try:
                    v = float(field)
                except ValueError:
                    continue
                if v == 0.0 and str(field).strip().startswith("-"):
                    raise SystemExit(f"{name}: -0 in {line!r}")
        body = "".join(f'    "{line}\\n"\n' for line in text.splitlines())
        parts.append(f"static const char {sym}[] =\n{body};\n")
    new = "// generated by gen_tables.py from spectral_amd/data -- do not edit\n" + "".join(parts)
    if not out.exists() or out.read_text() != new:
        out.write_text(new)


if __name__ == "__main__":
    main()

"""Writes reference_kats.json: the known-answer data of SmaxKernels' own SpMV /
SpMM unit tests (examples/tests/*.hpp), as plain numbers.

Run once in a container that holds the reference sources:

    python tests/golden/extract_reference_kats.py /root/reference

For every test header it records each array initialiser
(`name = new T[...]{...}`), each scalar metadata assignment
(`name->field = value;`), each constructor argument list
(`new SCSMatrix<...>(C, sigma)`, `new BCRSMatrix<...>(...)`,
`new CRSMatrix<...>(...)`) and each `const int|bool name = value;`, with the
source line it came from.  Only data is kept; no source text is stored.
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

TESTS = [
    "spmv_test.hpp",
    "spmv_scs_1_1_test_1.hpp", "spmv_scs_1_4_test_1.hpp",
    "spmv_scs_2_1_test_1.hpp", "spmv_scs_2_4_test_1.hpp",
    "spmv_scs_4_1_test_1.hpp", "spmv_scs_4_4_test_1.hpp",
    "spmv_scs_1_1_test_2.hpp", "spmv_scs_2_1_test_2.hpp",
    "spmv_scs_4_1_test_2.hpp",
    "spmv_bcrs_1_1_test_1.hpp", "spmv_bcrs_1_1_test_2.hpp",
    "spmv_bcrs_1_1_test_3.hpp",
    "spmm_test.hpp", "spmm_rowmaj_test.hpp",
]

ARRAY = re.compile(r"([\w>\-]+)\s*=\s*\n?\s*new\s+\w+\s*\[[^\]]*\]\s*\{([^}]*)\}",
                   re.S)
SCALAR = re.compile(r"(\w+->\w+)\s*=\s*([-+0-9.eE]+)\s*;")
CTOR = re.compile(r"\*\s*(\w+)\s*=\s*new\s+(\w+)<[^>]*>\s*\(([^)]*)\)")
CONST = re.compile(r"const\s+(?:int|bool)\s+(\w+)\s*=\s*(\w+)\s*;")
NVEC = re.compile(r"int\s+(n_vectors)\s*=\s*(\d+)\s*;")


def number(tok: str):
    tok = tok.strip()
    if tok in ("true", "false"):
        return tok == "true"
    return float(tok) if any(c in tok for c in ".eE") else int(tok)


def line_of(text: str, pos: int) -> int:
    return text.count("\n", 0, pos) + 1


def extract(path: Path) -> dict:
    text = path.read_text()
    out = {"arrays": {}, "scalars": {}, "ctors": {}, "consts": {}}
    for m in ARRAY.finditer(text):
        vals = [number(t) for t in m.group(2).split(",") if t.strip()]
        out["arrays"][m.group(1)] = {"line": line_of(text, m.start()),
                                     "values": vals}
    for m in SCALAR.finditer(text):
        out["scalars"][m.group(1)] = {"line": line_of(text, m.start()),
                                      "value": number(m.group(2))}
    for m in CTOR.finditer(text):
        args = [number(t) for t in m.group(3).split(",") if t.strip()]
        out["ctors"][m.group(1)] = {"line": line_of(text, m.start()),
                                    "type": m.group(2), "args": args}
    for rx in (CONST, NVEC):
        for m in rx.finditer(text):
            out["consts"][m.group(1)] = {"line": line_of(text, m.start()),
                                         "value": number(m.group(2))}
    return out


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    tests = ref / "examples" / "tests"
    data = {}
    for name in TESTS:
        data[name] = extract(tests / name)
        data[name]["source"] = f"examples/tests/{name}"
    out = Path(__file__).resolve().parent / "reference_kats.json"
    out.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {out} ({len(data)} tests)")


if __name__ == "__main__":
    main()

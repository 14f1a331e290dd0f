"""Writes sort_digests.json: SHA-256 digests of the SELL-C-sigma permutation
produced by the system libstdc++'s real std::sort (sort_pin.cpp) on the
Lomax-row-length workload, at test and at full benchmark size.

    python tests/golden/make_sort_digests.py

The tie order of std::sort decides perm (and with it col/val) for
sigma >= 17; these digests pin it where no reference test does.
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from smaxkernels_amd import matgen  # noqa: E402

SIZES = (30000, 4_000_000)


def main() -> None:
    tmp = Path(tempfile.mkdtemp())
    exe = tmp / "sort_pin"
    subprocess.run(["g++", "-O2", "-std=c++17", str(HERE / "sort_pin.cpp"),
                    "-o", str(exe)], check=True)
    cases = []
    for N in SIZES:
        lib = matgen._load()
        rp = np.empty(N + 1, np.int32)
        nnz = lib.smax_gen_lomax_rowptr(N, matgen.SEED_MATRIX, matgen.LOMAX_LAMBDA,
                                        matgen.LOMAX_CAP, N // 8, rp.ctypes.data)
        for C in (32, 64):
            npad = -(-N // C) * C
            cnt = np.zeros(npad, np.int32)
            cnt[:N] = np.diff(rp)
            for sigma in (1, 17, 64, 4096, npad):
                cnt.tofile(tmp / "c.bin")
                subprocess.run([str(exe), "sort", str(sigma), str(tmp / "c.bin"),
                                str(tmp / "o.bin")], check=True)
                order = np.fromfile(tmp / "o.bin", np.int32)
                slot = np.empty(npad, np.int64)
                slot[order] = np.arange(npad)
                perm = slot[:N].astype(np.int32)
                srt = cnt[order]
                chunk_len = srt.reshape(-1, C).max(axis=1).astype(np.int32)
                cases.append({
                    "N": N, "nnz": int(nnz), "C": C, "sigma": int(sigma),
                    "n_elements": int(chunk_len.astype(np.int64).sum() * C),
                    "perm_sha256": hashlib.sha256(perm.tobytes()).hexdigest(),
                    "chunk_lengths_sha256":
                        hashlib.sha256(chunk_len.tobytes()).hexdigest(),
                })
                print(cases[-1], flush=True)
    out = {"generator": {"lambda": matgen.LOMAX_LAMBDA, "cap": matgen.LOMAX_CAP,
                         "seed": matgen.SEED_MATRIX, "half": "N//8"},
           "sort": "libstdc++ std::sort, comparator a.count > b.count",
           "cases": cases}
    (HERE / "sort_digests.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()

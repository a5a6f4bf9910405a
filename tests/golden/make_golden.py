"""Writes the committed golden fixtures of tests/golden/.

* ref_product.npz — seeded skew matrices (n = 2..32) and the product-form
  Pfaffian returned by the reference's own CPU code (src/pfaffian_cpu.cpp),
  compiled from its source by oracle/Makefile into oracle/_ref.  These pin
  the oracle's stale-pivot mode bit for bit without the reference tree.
* mp_slog.npz — seeded real and complex skew matrices (n = 6..48) with
  log|Pf| and sign/phase from an mpmath Parlett-Reid at 50 digits.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def ref_product():
    rng = np.random.default_rng(20261203)
    mats, sizes, vals = [], [], []
    for n in range(2, 33, 2):
        for _ in range(3):
            A = oracle.skew_gaussian(rng, n) * 2.0
            mats.append(A.reshape(-1))
            sizes.append(n)
            vals.append(oracle.ref_pfaffian(A))
    np.savez_compressed(os.path.join(OUT, "ref_product.npz"), data=np.concatenate(mats),
                        sizes=np.array(sizes, np.int32), pf=np.array(vals))


def mp_slog():
    rng = np.random.default_rng(777)
    out = {}
    for kind in ("real", "complex"):
        mats, sizes, las, phs = [], [], [], []
        for n in (6, 8, 12, 16, 24, 32, 40, 48):
            A = oracle.skew_gaussian(rng, n, complex_=(kind == "complex"))
            la, ph, _ = oracle.pfaffian_mp(A, 50)
            mats.append(A.reshape(-1))
            sizes.append(n)
            las.append(la)
            phs.append(ph)
        out[f"{kind}_data"] = np.concatenate(mats)
        out[f"{kind}_sizes"] = np.array(sizes, np.int32)
        out[f"{kind}_log"] = np.array(las)
        out[f"{kind}_phase"] = np.array(phs)
    np.savez_compressed(os.path.join(OUT, "mp_slog.npz"), **out)


if __name__ == "__main__":
    if not oracle.ref_available():
        oracle.build()
    ref_product()
    mp_slog()
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))

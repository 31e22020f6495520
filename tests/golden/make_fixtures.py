"""Writes the golden fixtures of this directory.

1. wavefield-2d-so-{2,8}.npz: the reference's own golden wavefields, copied
   as data from its tests/reference_solution/wavefield-2d-so-{2,8}.npy
   (made by its tests/reference_solution/generator.py; loaded with
   allow_pickle=False).  Its 3D counterparts are not shipped with the
   reference.
2. ref_outputs.json: outputs of the reference C forward (its sequential
   `c` build: the OpenMP build adds overlapping sources with atomics in
   thread order, so it is not reproducible), compiled from its own sources
   into oracle/_ref (`make -C oracle ref`), on the reference's
   golden configurations (tests/test_solution.py:10-23 of the reference,
   3D included) and on kernel-level cases of tests/cases.py: SHA-256 of the
   returned u_full and shot record (bit identity), their shapes and
   float64 norms, and the shot record itself where it is small.

Run from the repository root with the reference at /root/reference:
    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, TESTS)
sys.path.insert(0, os.path.dirname(TESTS))

import cases  # noqa: E402
import harness  # noqa: E402

REFERENCE = os.environ.get("SIMWAVE_REFERENCE", "/root/reference")

# name -> builder; the reference's golden grid first, then kernel-level cases
GOLDEN = {
    "golden_2d_cd_so2": lambda: cases.golden_2d(2),
    "golden_2d_cd_so8": lambda: cases.golden_2d(8),
    "golden_2d_vd_so2": lambda: cases.golden_2d(2, density=True),
    "golden_3d_cd_so2": lambda: cases.golden_3d(2),
    "golden_3d_cd_so8": lambda: cases.golden_3d(8),
    "golden_3d_vd_so2": lambda: cases.golden_3d(2, density=True),
}
KERNEL = {
    "2d_cd_so20": dict(dim=2, space_order=20),
    "2d_vd_f64_so8": dict(dim=2, space_order=8, density=True, dtype=np.float64),
    "2d_stride3": dict(dim=2, stride=3, tf=0.08),
    "2d_src_many": dict(dim=2, src=[(100.0 + 3 * i, 50.0 + 5 * i) for i in range(70)],
                        window=2),
    "3d_vd_quirk": dict(dim=3, shape=(26, 30, 24), space_order=4, density=True,
                        tf=0.08),
    "3d_cd_f64_so8": dict(dim=3, n=26, space_order=8, dtype=np.float64, tf=0.08),
    "3d_bc_mixed": dict(dim=3, n=26, bc=("null_dirichlet", "null_neumann", "none",
                                         "null_neumann", "null_neumann",
                                         "null_dirichlet"), damping=0.0, tf=0.15),
    "3d_stride2": dict(dim=3, n=24, stride=2, tf=0.05),
    "3d_wide_bc": dict(dim=3, shape=(25, 29, 275), space_order=8, damping=0.0,
                       bc=("null_neumann", "null_dirichlet", "null_neumann",
                           "null_neumann", "null_neumann", "null_neumann"),
                       src=[(0.0, 100.0, 2700.0), (60.0, 140.0, 1300.0)], multi=True,
                       tf=0.05),
}


def digest(a):
    a = np.ascontiguousarray(a)
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "shape": list(a.shape), "dtype": str(a.dtype),
            "l2": float(np.linalg.norm(a.astype(np.float64))),
            "max_abs": float(np.abs(a).max()) if a.size else 0.0}


def setup_of(name):
    if name in GOLDEN:
        return GOLDEN[name]()
    return cases.make_case(**KERNEL[name])


def main():
    for so in (2, 8):
        src = os.path.join(REFERENCE, "tests", "reference_solution",
                           "wavefield-2d-so-%d.npy" % so)
        u = np.load(src, allow_pickle=False)
        np.savez_compressed(os.path.join(HERE, "wavefield-2d-so-%d.npz" % so), u=u)
        print("copied", src)
    if not harness.have("ref", omp=False):
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    out = {}
    for name in list(GOLDEN) + list(KERNEL):
        u, rec = harness.run(setup_of(name), "ref", omp=False)
        entry = {"u": digest(u), "rec": digest(rec)}
        if rec.size <= 4096:
            entry["rec_values"] = rec.astype(np.float64).ravel().tolist()
        out[name] = entry
        print(name, entry["u"]["shape"], entry["u"]["sha256"][:12])
    with open(os.path.join(HERE, "ref_outputs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

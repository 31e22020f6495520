"""Writes tests/golden/config_outputs.json: digests of the reference C
forward on the benchmark-size experiments of simwave_amd.workloads.

The reference's own cpu_openmp build (its wave.c files compiled from their
sources into oracle/_ref by `make -C oracle ref`, flags of its
compiler.py:49,59-66) runs each experiment in full in this container; every
experiment has one source, so its atomic source adds are in a fixed order
and the build is deterministic.  The GPU tests (tests/test_gpu_configs.py)
rebuild the same inputs, run the HIP library and compare:

* SHA-256 of the returned u_full and shot record (bit identity), with
  their shapes and float64 norms;
* SHA-256 of the extended velocity, damping mask and (where present)
  density, so that a mismatch is attributed to the inputs or the kernels.

    python tests/golden/make_config_fixtures.py [name ...]

Names: readme_2d, headline_3d_f32, headline_3d_f64, faulted_2d_vd_f32.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, TESTS)
sys.path.insert(0, ROOT)

import harness  # noqa: E402
import simwave_amd as sw  # noqa: E402
from simwave_amd import workloads  # noqa: E402
from simwave_amd.middleware import Middleware  # noqa: E402

OUT = os.path.join(HERE, "config_outputs.json")

EXPERIMENTS = {
    "readme_2d": lambda: workloads.readme_2d(np.float32),
    "headline_3d_f32": lambda: workloads.headline_3d(np.float32),
    "headline_3d_f64": lambda: workloads.headline_3d(np.float64),
    "faulted_2d_vd_f32": lambda: workloads.faulted_2d_vd(np.float32),
}


def digest(a):
    a = np.ascontiguousarray(a)
    return {"sha256": hashlib.sha256(a.view(np.uint8)).hexdigest(),
            "shape": list(a.shape), "dtype": str(a.dtype),
            "l2": float(np.linalg.norm(a.astype(np.float64).ravel()))}


def input_digests(solver):
    kw = Middleware.normalize(**solver.forward_arguments())
    out = {}
    for key in ("velocity_model", "density_model", "damping_mask"):
        if kw.get(key) is not None:
            out[key] = hashlib.sha256(np.ascontiguousarray(kw[key]).view(np.uint8)).hexdigest()
    out["timesteps"] = int(kw["end_timestep"])
    out["num_snapshots"] = int(kw["num_snapshots"])
    return out


def main(names):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        parts = EXPERIMENTS[name]()
        solver = sw.Solver(compiler=harness.LibraryCompiler("ref", omp=True), **parts)
        solver._middleware.verbose = False
        inputs = input_digests(solver)
        print(name, "steps", inputs["timesteps"], flush=True)
        t0 = time.time()
        u, rec = solver.forward()
        secs = time.time() - t0
        data[name] = {"u": digest(u), "rec": digest(rec), "inputs": inputs,
                      "source": "reference cpu_openmp build (oracle/_ref)",
                      "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
                      "seconds": round(secs, 1)}
        print(name, "done in %.0f s" % secs, data[name]["u"]["sha256"][:16],
              data[name]["rec"]["sha256"][:16], flush=True)
        del u, rec, solver, parts
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(EXPERIMENTS))

"""Writes the small .bin fixtures under tests/golden/ (run in the build
container; the GPU tests only read the committed files).

* cube-init.output00029.bin: the reference's own output file
  (output/cube-init.output00029.bin, v2.11, 1000 particles), copied as is.
* init-1e4-slab.bin: the particles of the reference's settled bed
  (output/init-1e4.output00199.bin, v2.11) with centres below z = 0.1 m,
  upgraded to v2.15 by sphere_amd's reader and written by its writer.
* shear1k.bin: a 1000-particle direct-shear cell (lognormal radii, median
  0.01 m, sigma 0.2) settled under gravity and consolidated at 10 kPa by the
  CPU oracle (sphere_amd.scenes.settled_shear_cell), seed 1105, then set up
  for shear (fixed bottom layer, top layer at 1e-3 m/s).

fixtures.json records the SHA-256 of each file.
"""

import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF = "/root/reference/output"


def sha(fn):
    return hashlib.sha256(open(fn, "rb").read()).hexdigest()


def main():
    import numpy
    import oracle
    from sphere_amd import scenes
    from sphere_amd.sim import sim

    out = {}
    src = os.path.join(REF, "cube-init.output00029.bin")
    dst = os.path.join(HERE, "cube-init.output00029.bin")
    shutil.copyfile(src, dst)
    out["cube-init.output00029.bin"] = sha(dst)

    s = sim()
    s.readbin(os.path.join(REF, "init-1e4.output00199.bin"), verbose=False)
    keep = numpy.nonzero(s.x[:, 2] < 0.1)[0]
    drop = numpy.setdiff1d(numpy.arange(int(s.np[0])), keep)
    s.deleteParticle(drop)
    s.sid = "init-1e4-slab"
    dst = os.path.join(HERE, "init-1e4-slab.bin")
    with open(dst, "wb") as fh:
        fh.write(s.tobytes())
    out["init-1e4-slab.bin"] = sha(dst)

    oracle.set_threads(os.cpu_count() or 1)
    s = scenes.settled_shear_cell(1000, 1105, open_stepper=oracle.Oracle)
    s.sid = "shear1k"
    dst = os.path.join(HERE, "shear1k.bin")
    with open(dst, "wb") as fh:
        fh.write(s.tobytes())
    out["shear1k.bin"] = sha(dst)

    with open(os.path.join(HERE, "fixtures.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

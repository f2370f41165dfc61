"""Convert the reference's level-set test results into fixtures.

Reads (data only, nothing is executed) from the reference checkout:
  Tests/LevelSetTests/{MergingCircles,SlottedDisk,VortexDeformation}/
      {FullDomain,NarrowBand}/
    input.st                       deck (mesh, time step, reinitialization)
    log.out                        number of steps and of reinitializations
    results/solution_0000.vtr      phi at t = 0 (PETSc VTK, appended raw FP64)
    results/solution_00NN.vtr      phi at the final time
and writes tests/golden/levelset_tests.npz: the final phi of each test in
full precision, and the SHA-256 of its t = 0 field (the full-domain tests
regenerate that field from the analytic initial condition and check the
digest; in the narrow-band tests it is the field after the band is built and
phi is cut off outside it).  Narrow-band keys carry the suffix "NB".  Run:
  python tests/golden/make_levelset_fixtures.py /root/reference
"""
import glob
import hashlib
import os
import re
import sys

import numpy as np

TESTS = ("MergingCircles", "SlottedDisk", "VortexDeformation")


def read_vtr(path, name="levelset0"):
    """Point data `name` of a rectilinear-grid file written piece by piece
    with appended raw data (UInt64 block headers), as a [k][j][i] array."""
    raw = open(path, "rb").read()
    head_end = raw.index(b"<AppendedData")
    head = raw[:head_end].decode()
    assert 'header_type="UInt64"' in head
    base = raw.index(b"_", head_end) + 1
    ext = list(map(int, re.search(r'WholeExtent="(\d+) (\d+) (\d+) (\d+) (\d+) (\d+)"', head).groups()))
    out = np.full((ext[5] - ext[4] + 1, ext[3] - ext[2] + 1, ext[1] - ext[0] + 1), np.nan)
    for pm in re.finditer(r'<Piece Extent="(\d+) (\d+) (\d+) (\d+) (\d+) (\d+)">(.*?)</Piece>', head, re.S):
        e = list(map(int, pm.groups()[:6]))
        off = base + int(re.search(r'Name="%s"[^>]*offset="(\d+)"' % name, pm.group(7)).group(1))
        nbytes = int(np.frombuffer(raw[off:off + 8], "<u8")[0])
        a = np.frombuffer(raw[off + 8:off + 8 + nbytes], "<f8")
        out[e[4]:e[5] + 1, e[2]:e[3] + 1, e[0]:e[1] + 1] = a.reshape(e[5] - e[4] + 1, e[3] - e[2] + 1,
                                                                    e[1] - e[0] + 1)
    assert not np.isnan(out).any()
    return out


def main(ref):
    out = {}
    for t in TESTS:
        for sub, suffix in (("FullDomain", ""), ("NarrowBand", "NB")):
            d = os.path.join(ref, "Tests", "LevelSetTests", t, sub)
            snaps = sorted(glob.glob(os.path.join(d, "results", "solution_*.vtr")))
            phi0 = read_vtr(snaps[0])
            phin = read_vtr(snaps[-1])
            log = open(os.path.join(d, "log.out")).read()
            steps = [int(m) for m in re.findall(r"^Step (\d+):", log, re.M)]
            key = t + suffix
            out[key + "_final"] = phin[0]
            out[key + "_initial_sha256"] = np.frombuffer(hashlib.sha256(phi0[0].tobytes()).digest(), np.uint8)
            out[key + "_steps"] = np.int64(steps[-1])
            out[key + "_reinit_calls"] = np.int64(len(re.findall(r"^- Reinitializing the level set", log, re.M)))
            print(key, phin.shape, "steps", steps[-1], "reinit", int(out[key + "_reinit_calls"]),
                  os.path.basename(snaps[-1]))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "levelset_tests.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
